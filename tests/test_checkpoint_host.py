"""Binary checkpoint files on the CPU: fgpe_checkpoint_info reads a file
written here from the layout documented in include/fgpe.h and
fortress_amd/csrc/fgpe_ckpt.hip (an independent writer of the format), checks
its FNV-1a checksum, and refuses corrupted, truncated, foreign and trailing
data with FGPE_EINVAL.  The device side (save, load, bit-identical resume) is
tests/test_gpu_checkpoint.py."""
import struct

import numpy as np
import pytest

from fortress_amd import fgpe


def fnv1a64(data):
    h = 0xcbf29ce484222325
    # FNV-1a 64 over every byte
    for b in data:
        h ^= b
        h = (h * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def make_file(path, dim=2, spin=1, precision=1, imag=1, n=(8, 4, 1), ii=1234, en_old=3.5, payload=None,
              magic=b"FGPECKP1", version=1):
    ms = 2 * spin + 1
    npts = n[0] * n[1] * n[2]
    if payload is None:
        rng = np.random.default_rng(0)
        payload = rng.standard_normal(npts * ms * 2).astype(np.float32 if precision == 1 else np.float64).tobytes()
    head = magic
    head += struct.pack("<6i", version, dim, spin, precision, imag, 1)
    head += struct.pack("<3i", *n)
    head += struct.pack("<3d", 0.1, 0.1, 1.0) + struct.pack("<d", 1e-3)
    head += struct.pack("<3d", 0.5, 0.5, 0.0) + struct.pack("<3d", 10.0, -1.0, 0.5)
    head += struct.pack("<3d", 1.0, 1.0, 20.0) + struct.pack("<d", 0.0)
    head += struct.pack("<q", ii) + struct.pack("<d", en_old) + struct.pack("<6d", *range(6))
    head += struct.pack("<8d", *([0.2] * 8)) + struct.pack("<2d", 20.0, 20.0)
    head += struct.pack("<q", len(payload))
    body = head + payload
    with open(path, "wb") as f:
        f.write(body + struct.pack("<Q", fnv1a64(body)))
    return len(payload)


def test_info_reads_header_and_verifies_checksum(tmp_path):
    p = str(tmp_path / "a.ckpt")
    nbytes = make_file(p)
    info = fgpe.checkpoint_info(p)
    assert info == dict(version=1, dim=2, spin=1, precision=1, imag=1, n=(8, 4, 1), dt=1e-3, ii=1234, en_old=3.5,
                        payload_bytes=nbytes)


@pytest.mark.parametrize("damage", ["payload_bit", "header_bit", "truncated", "trailing", "magic", "version"])
def test_info_refuses_damaged_files(tmp_path, damage):
    p = str(tmp_path / "b.ckpt")
    if damage == "magic":
        make_file(p, magic=b"FGPECKP0")
    elif damage == "version":
        make_file(p, version=2)
    else:
        make_file(p)
        raw = bytearray(open(p, "rb").read())
        if damage == "payload_bit":
            raw[-20] ^= 0x01
        elif damage == "header_bit":
            raw[40] ^= 0x80
        elif damage == "truncated":
            raw = raw[:-3]
        elif damage == "trailing":
            raw += b"\0"
        open(p, "wb").write(bytes(raw))
    with pytest.raises(fgpe.FgpeError):
        fgpe.checkpoint_info(p)


def test_info_missing_file(tmp_path):
    with pytest.raises(fgpe.FgpeError, match="cannot open"):
        fgpe.checkpoint_info(str(tmp_path / "none.ckpt"))
