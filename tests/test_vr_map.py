"""The channel -> VR plane map of the exchange (fock.vr_plane_map) on the CPU:
identical planes share the first one's index, planes differing in a single
bit never do, and a VR with all planes distinct needs no map."""

import numpy as np
import torch

from jax_hf_amd import synthetic as syn
from jax_hf_amd.fock import vr_plane_map


def _planes(V):
    nb = V.shape[-1]
    return torch.from_numpy(np.ascontiguousarray(np.moveaxis(V.reshape(V.shape[0], V.shape[1], nb * nb), -1, 0)))


def test_layer_resolved_v_maps_to_its_layer_pairs():
    nk, nb = 6, 8
    for nl, distinct in ((2, 2), (4, 4)):   # V depends on |l - l'| only
        V = syn.coulomb_layer(nk, nk, nb, nl).astype(np.complex128)
        m = vr_plane_map(_planes(V))
        assert m is not None
        m = m.numpy()
        assert len(np.unique(m)) == distinct
        lay = syn.layer_of_band(nb, nl)
        for c in range(nb * nb):
            i, j = divmod(c, nb)
            r = int(m[c])
            ri, rj = divmod(r, nb)
            assert r <= c and abs(lay[ri] - lay[rj]) == abs(lay[i] - lay[j])
            assert np.array_equal(V[..., i, j], V[..., ri, rj])
        # every channel's first equal plane, in index order
        assert m[0] == 0


def test_planes_differing_in_one_bit_stay_apart():
    rng = np.random.default_rng(0)
    X = rng.normal(size=(4, 5, 7)) + 1j * rng.normal(size=(4, 5, 7))
    X[2] = X[0]
    X[3] = X[0]
    X[3, 2, 3] = np.nextafter(X[3, 2, 3].real, np.inf) + 1j * X[3, 2, 3].imag
    m = vr_plane_map(torch.from_numpy(X)).numpy()
    assert list(m) == [0, 1, 0, 3]
    assert vr_plane_map(torch.from_numpy(rng.normal(size=(5, 3, 3)) + 0j)) is None
    c64 = torch.from_numpy(np.stack([X[0], X[0]]).astype(np.complex64))
    assert list(vr_plane_map(c64).numpy()) == [0, 0]
