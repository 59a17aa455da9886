"""Why the dense configuration's S is compared relative to the magnitude it is summed from (CPU, oracle only).

The dense configuration scales each camera's Jacobian blocks by 10^U(-2, 2), so blocks joining a 1e2-scaled camera
with a 1e-2-scaled one are sums whose terms cancel to far below their size. Two orders of the same exact
computation then disagree by far more than 1e-12 relative per block: here the reference's association
(E^-1 C_1)' C_2 (the oracle, src/schur_reduction.jl:349) against C_1' (E^-1 C_2) summed landmark by landmark in
numpy, on the same E^-1. Both stay within a few eps of the magnitude |R| + sum |C_1|' |E^-1| |C_2| per block, the
standard bound of any summation order; the GPU tests hold the device to 1e-12 of that magnitude.
"""
import numpy as np

import oracle
from helpers import dense_magnitude_blocks, dense_upper_blocks, max_block_rel, max_block_rel_to


def test_association_orders_differ_beyond_relative_bound(blsamd):
    p = blsamd.generators.config3(scale=0.001)   # 256 cameras, 1000 landmarks of k ~ U{2..64} cameras
    st = p.structure
    orc = oracle.schur_reduce_problem(p)
    assert orc["ok"]
    nr = st.nr
    S0 = np.zeros(nr * nr)
    oracle.reset_schur_dense(S0, np.asfortranarray(p.R).T.reshape(-1).astype(np.float64), p.scaling_r, p.lam)
    S = S0.reshape(nr, nr, order="F").copy()
    for e in range(st.neblocks):
        de, eo = st.e_dims[e], st.eoffsets[e] - 1
        Ei = orc["E_inv"][eo:eo + de * de].reshape(de, de, order="F")
        cols, idx = [], []
        for c in range(st.eliminated_ptr[e] - 1, st.eliminated_ptr[e + 1] - 1):
            r = st.entry_retained[c] - 1
            dr, co = st.r_dims[r], st.coffsets[c] - 1
            cols.append(p.C_data[co:co + de * dr].reshape(dr, de).T)   # d_e x d_r
            idx.extend(range(st.boffsets_r[r] - 1, st.boffsets_r[r] - 1 + dr))
        Ce = np.hstack(cols)
        S[np.ix_(idx, idx)] -= Ce.T @ (Ei @ Ce)
    other = dense_upper_blocks(st, S.reshape(-1, order="F"))
    ref = dense_upper_blocks(st, orc["S"])
    rel = max_block_rel(other, ref)
    to_mag = max_block_rel_to(other, ref, dense_magnitude_blocks(p))
    # the per-block relative difference of two exact-arithmetic-equal orders exceeds 1e-12 (4.5e-11 here) ...
    assert rel > 1e-11
    # ... while both are within a few eps of the summands' magnitude (2.2e-16 here)
    assert to_mag < 1e-14
