"""Large copies between pageable host memory and the device go through the
context's pinned double buffer in 4 MiB chunks (2 MiB and more); smaller ones
take the runtime's path.  Round trips at sizes on both sides of those limits,
and of chunk boundaries, return the bytes unchanged, and the staged image
upload feeds the same gradient cache as the unstaged one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIB = 1 << 20


@pytest.mark.parametrize("nbytes", [1, 2 * MIB - 1, 2 * MIB, 4 * MIB, 4 * MIB + 1, 8 * MIB - 3, 12 * MIB + 5,
                                    21 * MIB + 7])
def test_buffer_round_trip(pkg, ctx, nbytes):
    rng = np.random.default_rng(nbytes)
    host = rng.integers(0, 256, nbytes, dtype=np.uint8)
    buf = ctx.alloc(nbytes)
    try:
        buf.upload(host)
        back = buf.download(np.uint8)
        assert back.tobytes() == host.tobytes()
        # a second round trip through the same slots, with other contents
        host2 = host[::-1].copy()
        buf.upload(host2)
        assert buf.download(np.uint8).tobytes() == host2.tobytes()
    finally:
        buf.free()


@pytest.mark.parametrize("shape", [(700, 700), (1024, 1024), (1300, 1900)])
def test_image_upload_and_gradient(pkg, ctx, shape):
    """Images of 1.9, 4 and 9.4 MiB: the device copy equals the host array
    and the gradient cache of the upload equals the one built from the
    device copy of the same data (uploaded through the unstaged path in
    blocks of 64 rows)."""
    H, W = shape
    img = np.random.default_rng(W).random((H, W), dtype=np.float32) * 255
    dimg = ctx.upload_image(img)
    assert ctx.download_image(dimg).tobytes() == img.tobytes()
    # the same pixels written in blocks of 64 rows (each copy below the staging limit)
    buf, blk = ctx.alloc(4 * W * H), ctx.alloc(4 * W * 64)
    for y in range(0, H, 64):
        rows = np.ascontiguousarray(img[y:y + 64])
        assert rows.nbytes < 2 * MIB
        blk.upload(rows)
        ctx.check(ctx.lib.dic_buffer_copy(ctx.handle, buf.ptr + 4 * W * y, blk.ptr, rows.nbytes))
    ctx.synchronize()
    blk.free()
    other = ctx.wrap_image(buf.ptr, W, H)
    g1 = ctx.ref_gradient(dimg, pkg.GRADIENT_CENTRAL, 5, 3, pkg.PRECISION_FP64).download()
    g2 = ctx.ref_gradient(other, pkg.GRADIENT_CENTRAL, 5, 3, pkg.PRECISION_FP64).download()
    assert all(np.array_equal(a, b) for a, b in zip(g1, g2))
    buf.free()
