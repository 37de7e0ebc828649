"""Write tests/golden/ref_vectors.npz from the reference's own code.

Runs in a container that has the reference tree: builds oracle/_ref
(oracle/Makefile, target `ref`) and records input/output vectors of the
reference functions that compile on their own (Halton::sample, hash_u32,
SIMD::invert_44_matrix, spectrum_xyz_to_p, Sobol::sample).  The vectors are data; the
tests compare the oracle and the GPU library against them on any machine.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import ora  # noqa: E402


def main():
    ora.build(with_ref=True)
    R = ora.ref()
    rng = np.random.default_rng(20261202)
    # Halton: every dimension, edge indices plus random ones
    dims, idxs = [], []
    edge = np.array([0, 1, 2, 3, 242, 243, 244, 59048, 59049, 2**24 - 1, 2**24, 2**24 + 1, 2**31 - 1, 2**31,
                     2**32 - 2, 2**32 - 1], np.uint64)
    for d in range(256):
        ix = np.concatenate([edge, rng.integers(0, 2**32, 48, dtype=np.uint64)])
        dims.append(np.full(ix.size, d, np.uint32))
        idxs.append(ix.astype(np.uint32))
    hd = np.concatenate(dims)
    hi = np.concatenate(idxs)
    hv = np.array([R.psyref_halton(int(d), int(i)) for d, i in zip(hd, hi)], np.float32)
    # hash_u32
    hn = rng.integers(0, 2**32, 4096, dtype=np.uint64).astype(np.uint32)
    hs = rng.integers(0, 2**32, 4096, dtype=np.uint64).astype(np.uint32)
    hh = np.array([R.psyref_hash_u32(int(a), int(b)) for a, b in zip(hn, hs)], np.uint32)
    # matrix inverse: random, identity, translations, rotations
    mats = [np.eye(4, dtype=np.float32)]
    for _ in range(255):
        mats.append(rng.normal(size=(4, 4)).astype(np.float32))
    mats = np.stack(mats)
    invs = []
    for m in mats:
        buf = m.reshape(16).copy()
        R.psyref_invert44(buf.ctypes.data_as(__import__("ctypes").c_void_p))
        invs.append(buf.copy())
    invs = np.stack(invs).reshape(-1, 4, 4)
    # spectrum for grey input (the 5 nm knots of 380..700 and midpoints)
    lam = np.arange(380.0, 700.01, 2.5, dtype=np.float32)
    xyz = np.ones(3, np.float32)
    spec = np.array([R.psyref_spectrum_xyz_to_p(float(l), xyz.ctypes.data_as(__import__("ctypes").c_void_p))
                     for l in lam], np.float32)
    # Sobol::sample: every 16th dimension plus the ends, edge and random
    # 64-bit indices (bits 52+ reach into the next dimension's columns, so
    # the last dimension only takes indices below 2^52), random scrambles
    sdims, sidx, sscr = [], [], []
    sedge = [0, 1, 2, 3, 2**20, 2**32 - 1, 2**32, 2**52 - 1]
    nd = R.psyref_sobol_dims()
    for d in sorted(set(list(range(0, nd, 16)) + [1, 2, nd - 2, nd - 1])):
        top = 64 if d < nd - 1 else 52
        ix = sedge + [2**52, 2**63, 2**64 - 1] if top == 64 else list(sedge)
        ix += [int(v) for v in rng.integers(0, 2**top - 1, 24, dtype=np.uint64, endpoint=True)]
        sdims += [d] * len(ix)
        sidx += ix
        sscr += [0 if k % 3 == 0 else int(v) for k, v in
                 enumerate(rng.integers(0, 2**64 - 1, len(ix), dtype=np.uint64, endpoint=True))]
    sdims = np.array(sdims, np.uint32)
    sidx = np.array(sidx, np.uint64)
    sscr = np.array(sscr, np.uint64)
    sval = np.array([R.psyref_sobol(int(d), int(i), int(c)) for d, i, c in zip(sdims, sidx, sscr)], np.float64)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.npz")
    np.savez_compressed(out, halton_dim=hd, halton_index=hi, halton_value=hv, hash_n=hn, hash_seed=hs,
                        hash_value=hh, inv_in=mats, inv_out=invs, spectrum_lambda=lam, spectrum_grey=spec, sobol_dim=sdims,
                        sobol_index=sidx, sobol_scramble=sscr, sobol_value=sval)
    print("wrote", out)


if __name__ == "__main__":
    main()
