"""The host generator of the unsplit benchmark's oct batches: chunked draws
from one numpy default_rng stream, regenerable without torch, and the
first chunk alone gives the first batches (the CPU baseline's sample)."""
import numpy as np

from ramses_amd import ics


def test_oct_batches_host_chunks_and_first_chunk():
    uin, dt = ics.oct_batches_host(4 * 32 * 3, nv=32, seed=7, device="cpu", chunk=4)
    assert uin.shape == (12, 5, 6, 6, 6, 32)
    rng = np.random.default_rng(7)
    parts = [ics.oct_batches_host_chunk(rng, 4) for _ in range(3)]
    assert np.array_equal(uin.numpy(), np.concatenate([p[0] for p in parts]))
    assert dt == min(p[1] for p in parts)
    first, _ = ics.oct_batches_host_chunk(np.random.default_rng(7), 4)
    assert np.array_equal(uin.numpy()[:4], first)
    # conserved state of the drawn primitives: rho > 0 and E >= kinetic energy
    rho, e = first[:, 0], first[:, 4]
    kin = 0.5 * (first[:, 1] ** 2 + first[:, 2] ** 2 + first[:, 3] ** 2) / rho
    assert np.all(rho > 0) and np.all(e > kin)
