"""Host-side behaviour of the batch API while its device path is disabled."""
import pytest

from fortress_amd import fgpe
from fortress_amd.abi import FGPE_FP32, make_input


def test_batch_create_is_refused_with_a_clear_error():
    inp = make_input(2, 1, (64, 64), (0.1, 0.1), 1e-3, 87.0, (101.8, 100.4, 0.0), 10000, (20.0, 20.0, 400.0),
                     (0.5, 0.5), switch_im=1)
    p, _ = fgpe.derive(inp, FGPE_FP32)
    with pytest.raises(fgpe.FgpeError, match="disabled"):
        fgpe.Batch([p])
