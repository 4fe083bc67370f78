"""pirg_inner_product_fft_noalloc (device memory, static working array) against
pirg_inner_product_fft (host memory, allocating) through the C ABI.

Two calls in a row on one GPU: the ExternalProduct shape at d = 4096 (six
terms, 19-bit digits times 52-bit values, chunk 19,
crypto/external_product.cc:100-102), then a smaller signed FinalizeMatrixPack
shape (crypto/lwes_to_rlwe.cc:60-74), so the second call runs over what the
first left in the array. Both must equal the allocating entry point bit for
bit, and the oracle's InnerProductFft."""
import numpy as np
import pytest

import inspire_amd as ia
import oracle as O
from inspire_amd import hiprt

pytestmark = pytest.mark.gpu


def _run_noalloc(u, v, u_bits, v_bits, chunk, signed, stream=0):
    n, d = u.shape
    with hiprt.DeviceBuffer(u.nbytes) as du, hiprt.DeviceBuffer(v.nbytes) as dv, \
            hiprt.DeviceBuffer(d * 8) as do:
        du.upload(u)
        dv.upload(v)
        do.fill(0xA5)
        ia.inner_product_fft_noalloc(du.addr, dv.addr, n, d, u_bits, v_bits, chunk, signed,
                                     do.addr, stream)
        hiprt.synchronize()
        return do.download(np.uint64, d)


def test_noalloc_matches_allocating_entry_point_twice():
    rng = np.random.default_rng(2024)
    # Call 1: ExternalProduct at d = 4096.
    d = 4096
    u = rng.integers(0, 1 << 19, size=(6, d), dtype=np.uint64)
    v = rng.integers(0, 1 << 52, size=(6, d), dtype=np.uint64)
    print("noalloc call 1: n=6 d=4096 19x52 chunk 19", flush=True)
    got = _run_noalloc(u, v, 19, 52, 19, False)
    want = ia.inner_product_fft(u, v, 19, 52, 19)
    assert np.array_equal(got, want)
    ref, _ = O.inner_product_fft(u, v, 19, 52, 19)
    assert np.array_equal(got, ref)
    # Call 2, smaller: signed 19-bit digits (FinalizeMatrixPack) at d = 1024.
    d = 1024
    u = rng.integers(0, 1 << 64, size=(2, d), dtype=np.uint64)
    v = rng.integers(0, 1 << 52, size=(2, d), dtype=np.uint64)
    print("noalloc call 2: n=2 d=1024 signed 19x52 chunk 20", flush=True)
    got = _run_noalloc(u, v, 19, 52, 20, True)
    want = ia.inner_product_fft(u, v, 19, 52, 20, u_signed=True)
    assert np.array_equal(got, want)
    ref, _ = O.inner_product_fft(u, v, 19, 52, 20, u_signed=True)
    assert np.array_equal(got, ref)


def test_noalloc_rejects_sizes_beyond_its_array():
    # 40 terms of four u and four v chunks at d = 4096: 320 transforms of
    # 128 KiB, beyond the 32 MiB array. Rejected before any device work.
    with pytest.raises(ia.PirError) as e:
        ia.inner_product_fft_noalloc(1, 1, 40, 4096, 64, 52, 16, False, 1)
    assert e.value.code == ia.INVALID_ARGUMENT
    assert "no-allocation form" in e.value.message


def test_noalloc_argument_errors_match_reference_messages():
    with pytest.raises(ia.PirError, match="Polynomial vectors cannot be empty"):
        ia.inner_product_fft_noalloc(1, 1, 0, 1024, 19, 52, 20, False, 1)
    with pytest.raises(ia.PirError, match="Potential precision overflow"):
        ia.inner_product_fft_noalloc(1, 1, 2, 2048, 64, 64, 21, False, 1)
    with pytest.raises(ia.PirError, match="must be device pointers"):
        ia.inner_product_fft_noalloc(0, 1, 2, 1024, 19, 52, 20, False, 1)
