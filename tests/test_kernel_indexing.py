"""CPU checks of the index arithmetic the HIP kernels use, before any launch:
the Dslash neighbour formula (k_dslash in openstaple_amd/csrc/dslash.hip) must
reproduce the oracle's neighbour tables and stay in [0, sizeh), and the
XCD-aware workgroup remap must be a bijection of [0, G)."""
import itertools

import numpy as np
import pytest

from oracle import Oracle


def kernel_neighbours(nd, eo):
    """Vectorised restatement of the k_dslash index block for all threads."""
    nd0, nd1, nd2, nd3 = nd
    nd0h = nd0 // 2
    sizeh = nd0 * nd1 * nd2 * nd3 // 2
    idx = np.arange(sizeh, dtype=np.int64)
    t = idx // nd0h
    hd0 = idx - t * nd0h
    t2 = t // nd1
    d1 = t - t2 * nd1
    d3 = t2 // nd2
    d2 = t2 - d3 * nd2
    d0 = 2 * hd0 + ((d1 + d2 + d3 + eo) & 1)
    row = idx - hd0
    s1, s2 = nd0h, nd0h * nd1
    s3 = s2 * nd2
    ip0 = row + np.where(d0 + 1 == nd0, 0, (d0 + 1) >> 1)
    im0 = row + np.where(d0 == 0, (nd0 - 1) >> 1, (d0 - 1) >> 1)
    ip1 = idx + np.where(d1 == nd1 - 1, -(nd1 - 1) * s1, s1)
    im1 = idx + np.where(d1 == 0, (nd1 - 1) * s1, -s1)
    ip2 = idx + np.where(d2 == nd2 - 1, -(nd2 - 1) * s2, s2)
    im2 = idx + np.where(d2 == 0, (nd2 - 1) * s2, -s2)
    ip3 = idx + np.where(d3 == nd3 - 1, -(nd3 - 1) * s3, s3)
    im3 = idx + np.where(d3 == 0, (nd3 - 1) * s3, -s3)
    return np.stack([ip0, ip1, ip2, ip3], 1), np.stack([im0, im1, im2, im3], 1), sizeh


@pytest.mark.parametrize("shape", [(4, 4, 4, 4), (4, 6, 8, 10), (16, 4, 6, 2), (2, 2, 2, 2), (32, 32, 32, 32),
                                   (24, 24, 24, 8), (48, 48, 48, 12)])
@pytest.mark.parametrize("eo", [0, 1])
def test_dslash_neighbours_match_tables(shape, eo):
    ip, im, sizeh = kernel_neighbours(shape, eo)
    assert ip.min() >= 0 and ip.max() < sizeh and im.min() >= 0 and im.max() < sizeh
    if np.prod(shape) > 300000:
        return  # bounds only; the table comparison runs on the smaller shapes
    o = Oracle(*shape)
    nnp, nnm = o.nn_tables()
    par = 0 if eo == 0 else 1  # Deo writes even sites, Doe odd sites
    np.testing.assert_array_equal(ip, nnp[:, :, par])
    np.testing.assert_array_equal(im, nnm[:, :, par])


def xcd_block_id(i, G):
    q, r = G >> 3, G & 7
    x, k = i & 7, i >> 3
    return x * q + (x if x < r else r) + k


@pytest.mark.parametrize("G", list(range(1, 70)) + [256, 257, 1000, 2048, 4095, 4096, 5185])
def test_xcd_remap_is_a_bijection(G):
    ids = sorted(xcd_block_id(i, G) for i in range(G))
    assert ids == list(range(G))


def decode(idxh, parity, nd):
    """Restatement of `decode` in openstaple_amd/csrc/stout.hip."""
    nd0h = nd[0] >> 1
    t = idxh // nd0h
    hd0 = idxh - t * nd0h
    t2 = t // nd[1]
    d1 = t - t2 * nd[1]
    d3 = t2 // nd[2]
    d2 = t2 - d3 * nd[2]
    d0 = 2 * hd0 + ((d1 + d2 + d3 + parity) & 1)
    return [d0, d1, d2, d3]


def snum(d, nd):
    return (d[0] + nd[0] * (d[1] + nd[1] * (d[2] + nd[2] * d[3]))) >> 1


def shifted(d, mu, s, nu, t, nd):
    e = list(d)
    e[mu] = (e[mu] + s) % nd[mu]
    if t:
        e[nu] = (e[nu] + t) % nd[nu]
    return snum(e, nd)


@pytest.mark.parametrize("shape", [(4, 4, 4, 4), (4, 6, 8, 10), (16, 4, 6, 2)])
def test_link_kernel_site_decoding(shape):
    """decode/shifted of the per-link kernels (stout.hip, force.hip) agree with
    the neighbour tables and stay in bounds."""
    o = Oracle(*shape)
    nnp, nnm = o.nn_tables()
    sizeh = o.sizeh
    for parity in (0, 1):
        for i in range(sizeh):
            d = decode(i, parity, shape)
            assert snum(d, shape) == i and sum(d) % 2 == parity
            for mu in range(4):
                assert shifted(d, mu, 1, 0, 0, shape) == nnp[i, mu, parity]
                assert shifted(d, mu, -1, 0, 0, shape) == nnm[i, mu, parity]
                for nu in range(4):
                    if nu == mu:
                        continue
                    j = shifted(d, mu, 1, nu, -1, shape)
                    assert 0 <= j < sizeh
                    assert j == nnm[nnp[i, mu, parity], nu, 1 - parity]


def dslash_site_of_thread(G, nd3, sizeh, tmajor, block=256):
    """Restatement of the workgroup -> site mapping of k_dslash (dslash.hip)."""
    i = np.arange(G)
    q, r = G >> 3, G & 7
    x, k = i & 7, i >> 3
    blk = x * q + np.where(x < r, x, r) + k
    if tmajor:
        s3 = sizeh // nd3
        tile, tt = blk // nd3, blk % nd3
        base = tt * s3 + tile * block
    else:
        base = blk * block
    return (base[:, None] + np.arange(block)[None, :]).ravel()


@pytest.mark.parametrize("shape", [(32, 32, 32, 32), (24, 24, 24, 8), (32, 32, 32, 8), (48, 48, 48, 12),
                                   (16, 16, 16, 16), (4, 4, 4, 4), (4, 6, 8, 10)])
def test_dslash_workgroup_mapping_is_a_bijection(shape):
    sizeh = np.prod(shape) // 2
    s3 = sizeh // shape[3]
    tmajor = s3 % 256 == 0  # the launcher's condition
    G = (sizeh + 255) // 256
    idx = dslash_site_of_thread(G, shape[3], sizeh, tmajor)
    valid = idx[idx < sizeh]
    assert len(np.unique(valid)) == sizeh and valid.min() == 0
    if tmajor:
        assert idx.max() == sizeh - 1  # no thread beyond the lattice
