"""Ring-buffer segment copies (GpuRingbufCopyKernel; reference src_lib/RingbufCopyKernel.cu:23-240).

The oracle's restatement of CpuRingbufCopyKernel::apply is pinned by known answers worked out
by hand (phase arithmetic with wrap-around, in-order beams within an octuple); the GPU kernel
must reproduce the oracle byte for byte on random octuple sets built as the reference's
test_random builds them (source and destination zones apart, destination frame ranges of
different octuples disjoint), for 128- and 256-byte segments.
"""

import numpy as np
import pytest


def _segments(nseg, seg_bytes):
    """Ring of nseg segments; every byte of segment i holds (i + 1) mod 256."""
    return np.repeat((np.arange(nseg) + 1) % 256, seg_bytes).astype(np.uint8)


def _seg_ids(rb, seg_bytes):
    v = rb.reshape(-1, seg_bytes)
    assert np.all(v == v[:, :1])
    return [int(x) - 1 for x in v[:, 0]]


def test_oracle_known_answer_phases(orc):
    rb = _segments(12, 128)
    # src: offset 0, phase 3, 4 frames, 1 seg/frame; dst: offset 8, phase 0, 2 frames, 2 seg/frame.
    # iframe = 5, 2 beams: src frames (3+5)%4 = 0 then 1 -> segments 0, 1;
    #                       dst frames (0+5)%2 = 1 then 0 -> segments 10, 8.
    orc.ringbuf_copy(rb, [[0, 3, 4, 1, 8, 0, 2, 2]], 128, 2, 5)
    ids = _seg_ids(rb, 128)
    assert ids[10] == 0 and ids[8] == 1
    assert [ids[i] for i in range(12) if i not in (8, 10)] == [i for i in range(12) if i not in (8, 10)]


def test_oracle_known_answer_in_order(orc):
    # Beam 0 copies segment 0 -> 1, beam 1 copies segment 1 -> 2 (dst of beam 0 is src of
    # beam 1): in order, segment 2 ends up with segment 0's data.
    rb = _segments(4, 256)
    orc.ringbuf_copy(rb, [[0, 0, 3, 1, 1, 0, 3, 1]], 256, 2, 0)
    assert _seg_ids(rb, 256) == [0, 0, 0, 3]
    # Two octuples run in order: the second sees the first's result.
    rb = _segments(4, 128)
    orc.ringbuf_copy(rb, [[3, 0, 1, 1, 0, 0, 1, 1], [0, 0, 1, 1, 2, 0, 1, 1]], 128, 1, 7)
    assert _seg_ids(rb, 128) == [3, 1, 3, 3]


def test_oracle_rejects_out_of_range(orc):
    rb = _segments(4, 128)
    with pytest.raises(RuntimeError, match='outside'):
        orc.ringbuf_copy(rb, [[0, 0, 5, 1, 1, 0, 1, 1]], 128, 1, 0)


def test_create_rejects_bad_params():
    import pirate_amd as pa
    from pirate_amd._lib import PddError
    with pytest.raises(PddError, match='128 or 256'):
        pa.GpuRingbufCopyKernel(4, 2, 16, np.zeros((0, 8)), dtype='float32')
    with pytest.raises(PddError, match='multiple'):
        pa.GpuRingbufCopyKernel(5, 2, 32, np.zeros((0, 8)), dtype='float32')
    with pytest.raises(PddError, match='zero frame count'):
        pa.GpuRingbufCopyKernel(4, 2, 32, [[0, 0, 0, 1, 1, 0, 1, 1]], dtype='float32')
    k = pa.GpuRingbufCopyKernel(4, 2, 64, np.zeros((0, 8)), dtype='float16')   # no octuples: nothing on the GPU
    assert k.nbytes_per_segment == 128


def random_case(rng, seg_bytes):
    """(total_beams, beams_per_batch, ring nsegments, octuples) in the manner of the reference's
    test_random (RingbufCopyKernel.cu:322-400)."""
    nbatches, B = int(rng.integers(1, 5)), int(rng.integers(1, 5))
    zones = []
    base = 0
    for _ in range(int(rng.integers(2, 9))):
        nfr = int(rng.integers(B, 4 * B + 3))
        spf = int(rng.integers(1, 5))
        zones.append((base, nfr, spf))
        base += nfr * spf
    order = rng.permutation(len(zones))
    nsrc = max(1, len(zones) // 2)
    src = [zones[i] for i in order[:nsrc]]
    dst = [zones[i] for i in order[nsrc:]] or [src.pop()]
    rows = []
    for (db, dnf, dspf) in dst:
        for s in range(dspf):
            f = int(rng.integers(0, dnf))
            fend = f + dnf - B
            while f <= fend:
                if rng.random() < 0.5:
                    f += 1
                    continue
                sb, snf, sspf = src[int(rng.integers(0, len(src)))]
                rows.append([sb + int(rng.integers(0, sspf)), int(rng.integers(0, 10**6)), snf, sspf,
                             db + s, f + int(rng.integers(0, 1000)) * dnf, dnf, dspf])
                f += B
    return nbatches * B, B, base, np.array(rows, dtype=np.uint32).reshape(-1, 8)


@pytest.mark.gpu
@pytest.mark.parametrize('seg_bytes', [128, 256])
def test_gpu_matches_oracle(seg_bytes, orc):
    import torch
    import pirate_amd as pa
    rng = np.random.default_rng(seg_bytes)
    for trial in range(12):
        total, B, nseg, oct_ = random_case(rng, seg_bytes)
        k = pa.GpuRingbufCopyKernel(total, B, seg_bytes // 4, oct_, dtype='float32')
        init = rng.integers(0, 256, size=nseg * seg_bytes, dtype=np.uint8)
        ref = init.copy()
        dev = torch.from_numpy(init.copy()).cuda()
        for step in range(3):
            ichunk, ibatch = int(rng.integers(0, 1000)), int(rng.integers(0, total // B))
            orc.ringbuf_copy(ref, oct_, seg_bytes, B, ichunk * total + ibatch * B)
            k.launch(dev, ichunk, ibatch)
        torch.cuda.synchronize()
        got = dev.cpu().numpy()
        assert np.array_equal(got, ref), f'trial {trial}: {len(oct_)} octuples, B={B}'


@pytest.mark.gpu
def test_gpu_rejects_out_of_range_octuple():
    import torch
    import pirate_amd as pa
    k = pa.GpuRingbufCopyKernel(2, 1, 32, [[0, 0, 4, 1, 4, 0, 4, 1]], dtype='float32')   # segments 0..7
    small = torch.zeros(7 * 128, dtype=torch.uint8, device='cuda')
    with pytest.raises(RuntimeError, match='beyond'):
        k.launch(small, 0, 0)
    ok = torch.zeros(8 * 128, dtype=torch.uint8, device='cuda')
    k.launch(ok, 0, 0)
    torch.cuda.synchronize()
