"""Host replay of the device reduction (test helper).

oracle_ke_layout: the oracle's element matrices placed in a plan's Ke layout.
reduce_host:      the device reduction's arithmetic on the plan's maps, in
                  numpy (segment lengths/offsets, slots, A/D split, rhs).
"""
import numpy as np

from oracle import py_oracle


def oracle_ke_layout(prob, ke_base, ke_ld, n_ke):
    d = prob.desc()
    info = prob.info()
    blocks, n = [], 0
    for b in range(d.n_buckets):
        bd = d.buckets[b]
        for s in range(bd.n_slices):
            sd = bd.slices[s]
            ndf = 1 if sd.is_rhs else sd.n_dof
            blocks.append((bd.n_elem, sd.n_tf, ndf, n))
            n += bd.n_elem * sd.n_tf * ndf
    o = py_oracle.assemble(prob.desc_ptr(), prob.numeric(), info["ndofs"], info["nfree"], n_ke=n)
    ke = np.zeros(n_ke)
    for k, (ne, nt, ndf, off) in enumerate(blocks):
        src = o.ke[off:off + ne * nt * ndf].reshape(ne, nt, ndf)
        e, i, j = np.meshgrid(np.arange(ne), np.arange(nt), np.arange(ndf), indexing="ij")
        ke[ke_base[k] + (e * nt + i) * ke_ld[k] + j] = src
    return o, ke


def reduce_host(maps, ke, Arow, Drow, shape):
    nf = shape["nfree"]
    er = maps["erows"]
    lens = (er >> np.uint64(48)).astype(np.int64)
    offs = (er & np.uint64((1 << 48) - 1)).astype(np.int64)
    rows_of_seg = np.repeat(np.arange(nf), np.diff(maps["erow_ptr"]))
    starts = np.repeat(offs, lens)
    within = np.arange(lens.sum()) - np.repeat(np.cumsum(lens) - lens, lens)
    kidx = starts + within
    rows = np.repeat(rows_of_seg, lens)
    # slot blocks are padded to 4 entries per segment
    padded = (lens + 3) // 4 * 4
    sstart = np.repeat(np.cumsum(padded) - padded, lens)
    slots = maps["slots"][sstart + within]
    assert len(maps["slots"]) == padded.sum()
    lenA = np.diff(Arow)
    on = slots >= 0
    # folded orientations: bit 30 of a slot negates the entry
    neg = on & ((slots & (1 << 30)) != 0)
    slots = np.where(on, slots & ~(1 << 30), slots)
    vals = np.where(neg, -ke[kidx], ke[kidx])
    isA = on & (slots < lenA[rows])
    Aval = np.zeros(shape["nnzA"])
    Dval = np.zeros(shape["nnzD"])
    np.add.at(Aval, Arow[rows[isA]] + slots[isA], vals[isA])
    isD = on & ~isA
    np.add.at(Dval, Drow[rows[isD]] + slots[isD] - lenA[rows[isD]], vals[isD])
    rr = maps["rhs_erows"]
    roffs = (rr & np.uint64((1 << 48) - 1)).astype(np.int64)
    rneg = (rr >> np.uint64(63)) != 0
    rrows = np.repeat(np.arange(nf), np.diff(maps["rhs_ptr"]))
    rhs_free = np.zeros(nf)
    np.add.at(rhs_free, rrows, np.where(rneg, -ke[roffs], ke[roffs]))
    return Aval, Dval, rhs_free
