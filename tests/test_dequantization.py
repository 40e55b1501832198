"""int4 dequantization (reference: ReferenceDequantizationKernel::apply,
src_lib/GpuDequantizationKernel.cu:326-380). CPU tests pin the oracle against the closed
form on every nibble value, the sentinel, float16 subnormals/inf, and the nibble order;
the GPU test compares the HIP kernel with the oracle bit for bit (float32) and with the
oracle rounded to 16 bits (float16 / bfloat16)."""

import numpy as np
import pytest

from oracle import oracle as orc


def pack(vals):
    """int values in [-8, 7], last axis even -> uint8, low nibble = even index."""
    u = (np.asarray(vals, dtype=np.int64) & 0xf).astype(np.uint8)
    return (u[..., 0::2] | (u[..., 1::2] << 4)).astype(np.uint8)


def random_case(rng, B, F, T, special=False):
    vals = rng.integers(-8, 8, size=(B, F, T))
    sc = rng.uniform(-3, 3, size=(B, F, T // 256, 2)).astype(np.float16)
    if special:
        flat = sc.reshape(-1)
        flat[:8] = np.array([0.0, -0.0, 6.0e-8, -6.0e-8, 65504, 1e-5, np.inf, 2.0**-14], dtype=np.float16)
    return vals, sc


def test_oracle_closed_form():
    rng = np.random.default_rng(1)
    vals, sc = random_case(rng, 2, 3, 512, special=True)
    out = orc.dequantize(pack(vals), sc, 2, 3, 512)
    scale = np.repeat(sc[..., 0].astype(np.float32), 256, axis=-1)
    off = np.repeat(sc[..., 1].astype(np.float32), 256, axis=-1)
    with np.errstate(invalid='ignore', over='ignore'):
        want = np.where(vals == -8, np.float32(0), (scale * vals.astype(np.float32)).astype(np.float32) + off)
    assert np.array_equal(out, want.astype(np.float32), equal_nan=True)
    assert np.all(out[vals == -8] == 0)


def test_oracle_nibble_order():
    vals = np.zeros((1, 1, 256), dtype=np.int64)
    vals[0, 0, 0], vals[0, 0, 1] = 3, -2
    sc = np.zeros((1, 1, 1, 2), dtype=np.float16)
    sc[..., 0] = 1.0
    d = pack(vals)
    assert d[0, 0, 0] == (3 | (0xe << 4))
    out = orc.dequantize(d, sc, 1, 1, 256)
    assert out[0, 0, 0] == 3.0 and out[0, 0, 1] == -2.0 and np.all(out[0, 0, 2:] == 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', ['float32', 'float16', 'bfloat16'])
def test_gpu_dequantization_matches_oracle(dtype):
    import torch
    import pirate_amd as pa
    rng = np.random.default_rng(5)
    B, F, T = 3, 37, 1024
    vals, sc = random_case(rng, B, F, T, special=(dtype == 'float32'))
    d = pack(vals)
    want = orc.dequantize(d, sc, B, F, T)
    k = pa.GpuDequantizationKernel(dtype, B, F, T)
    tdt = getattr(torch, dtype)
    out = torch.full((B, F, T), 7.0, dtype=tdt, device='cuda')
    k.launch(out, torch.from_numpy(sc).cuda(), torch.from_numpy(d).cuda())
    torch.cuda.synchronize()
    got = out.cpu()
    if dtype == 'float32':
        assert np.array_equal(got.numpy(), want, equal_nan=True)
    else:
        ref = torch.from_numpy(want).to(tdt)
        assert torch.equal(got.view(torch.int16), ref.view(torch.int16))
