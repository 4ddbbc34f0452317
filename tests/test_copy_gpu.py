"""The 16-byte streaming copy (estl_stream_copy), the achievable-bandwidth reference the
benchmark reports beside the transforms: every byte arrives, in both load/store forms,
across partial last workgroups, and bad arguments are refused before any launch."""
import pytest

import estl_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbytes", [16, 16 * 1023, 16 * 1024, (1 << 20) + 48, 3 * (1 << 26) + 16 * 7])
@pytest.mark.parametrize("nontemporal", [False, True])
def test_stream_copy_is_exact(gpu, nbytes, nontemporal):
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device=gpu)
    estl_amd.fill_uniform(src, 80)
    dst = torch.full((nbytes // 4 + 8,), 7.0, dtype=torch.float32, device=gpu)  # 32 guard bytes
    estl_amd.stream_copy(src, dst[: nbytes // 4], nontemporal)
    torch.cuda.synchronize()
    assert torch.equal(dst[: nbytes // 4], src)
    assert torch.all(dst[nbytes // 4:] == 7.0)  # nothing written past the end


def test_stream_copy_rejects_bad_arguments(gpu):
    import torch
    x = torch.zeros(64, dtype=torch.float32, device=gpu)
    y = torch.zeros(64, dtype=torch.float32, device=gpu)
    lib = estl_amd.lib()
    assert lib.estl_stream_copy(x.data_ptr(), y.data_ptr(), 24, 1, None) == estl_amd.ERR_INVALID  # not 16 B
    assert lib.estl_stream_copy(x.data_ptr() + 4, y.data_ptr(), 16, 1, None) == estl_amd.ERR_INVALID  # misaligned
    assert lib.estl_stream_copy(None, y.data_ptr(), 16, 1, None) == estl_amd.ERR_INVALID
    assert lib.estl_stream_copy(x.data_ptr(), y.data_ptr(), 0, 0, None) == estl_amd.OK  # nothing to copy
    torch.cuda.synchronize()
    assert torch.all(y == 0)
