"""CPU: the operator-file reader refuses what is not a complete h2w operator
file before touching any device."""
import numpy as np
import pytest

import h2weak_amd as h2w


def test_refuses_foreign_and_truncated_files(tmp_path):
    p = tmp_path / "x.h2w"
    p.write_bytes(b"NOTANOPS" + bytes(100))
    with pytest.raises(h2w.H2WError) as e:
        h2w.Initialisation.load(str(p))
    assert e.value.code == h2w.H2W_EINVAL
    # right magic, header cut short
    p.write_bytes(b"H2WSKEL1" + np.array([3, 0, 1, 10], dtype=np.int32).tobytes())
    with pytest.raises(h2w.H2WError) as e:
        h2w.Initialisation.load(str(p))
    assert e.value.code == h2w.H2W_EINVAL
    # impossible header values
    hdr = np.array([7, 0, 1, 10, 64, 3, 1, 0], dtype=np.int32).tobytes() + np.int64(1000).tobytes()
    p.write_bytes(b"H2WSKEL1" + hdr + np.float64(1.0).tobytes())
    with pytest.raises(h2w.H2WError) as e:
        h2w.Initialisation.load(str(p))
    assert e.value.code == h2w.H2W_EINVAL
    with pytest.raises(h2w.H2WError):
        h2w.Initialisation.load(str(tmp_path / "missing.h2w"))
