"""The reference's head-file and restart-file formats: the oracle against the
reference's own files (byte for byte), the host-only restart reader of the
library against the oracle, and (GPU) the library's writers against the
reference's files.  Bit-exact throughout: the files are bytes."""
import numpy as np
import pytest

from msg123_amd import gpu
from oracle import io as oio
from tests.golden.cases import IO


def _fixture(golden):
    g = golden("io_box9")
    H = g["step_heads"]
    return g, float(g["len_scal"]), H, H.shape[1], g["head_file"].tobytes(), g["rest_file"].tobytes()


def test_oracle_head_file_matches_reference(golden):
    g, L, H, n, hf, _ = _fixture(golden)
    recs = oio.read_head_records(hf, n)
    assert [t for t, _ in recs] == [0.0, float(g["step_times"][-1])]
    h0 = gpu.BoxMesh(IO["io_box9"]()).arrays()["h0"]
    assert oio.head_record(0.0, h0, L) + oio.head_record(g["step_times"][-1], H[-1], L) == hf
    # the product is taken in FP64 and rounded once: a product in FP32 differs
    v32 = H[-1].astype(np.float32) * np.float32(L)
    assert not np.array_equal(v32, recs[1][1])


def test_oracle_restart_matches_reference(golden):
    g, L, H, n, _, rf = _fixture(golden)
    assert len(rf) == 8 * (n + 1)
    assert oio.restart_file(g["step_times"][-1], H[-1], L) == rf
    h, t = oio.read_restart(rf, n, L)
    assert t == float(g["step_times"][-1])
    # len_scal_inv is 1/len_scal rounded to FP32: the read-back is not the
    # written head, and an FP64 inverse would give other values
    assert np.abs(h - H[-1]).max() < 1e-8
    assert not np.array_equal(h, np.frombuffer(rf, "<f8")[1:] / L)


def test_library_read_restart_matches_oracle(golden, tmp_path):
    g, L, H, n, _, rf = _fixture(golden)
    path = tmp_path / "rest.bin"
    path.write_bytes(rf)
    h, t = gpu.read_restart(str(path), n, L)
    h_o, t_o = oio.read_restart(rf, n, L)
    assert np.array_equal(h, h_o) and t == t_o
    path.write_bytes(rf[:-8])  # one value short
    with pytest.raises(RuntimeError):
        gpu.read_restart(str(path), n, L)
    with pytest.raises(RuntimeError):
        gpu.read_restart(str(tmp_path / "missing.bin"), n, L)


@pytest.mark.gpu
def test_library_writers_match_reference(golden, tmp_path):
    g, L, H, n, hf, rf = _fixture(golden)
    box = gpu.BoxMesh(IO["io_box9"]())
    ctx = gpu.Context(box)
    try:
        head = str(tmp_path / "head.bin")
        ctx.set_state(box.arrays()["h0"], delt=1.0)
        ctx.write_head(head, 0.0, append=False)
        ctx.set_state(H[-1], delt=1.0)
        ctx.write_head(head, float(g["step_times"][-1]))
        assert open(head, "rb").read() == hf
        rest = str(tmp_path / "rest.bin")
        ctx.write_restart(rest, 1.0)   # replaced by the next call
        ctx.write_restart(rest, float(g["step_times"][-1]))
        assert open(rest, "rb").read() == rf
        # restart read -> state -> head record: len_scal_inv is 1/len_scal in
        # FP32 (1 + 1.5e-8 relative for len_scal = 10), so a restarted head
        # can land one float32 ulp away from the record written before
        h, t = gpu.read_restart(rest, n, L)
        ctx.set_state(h, delt=1.0)
        ctx.write_head(head, t, append=False)
        (t_a, a), = oio.read_head_records(open(head, "rb").read(), n)
        t_b, b = oio.read_head_records(hf, n)[1]
        assert t_a == t_b
        ulps = np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))
        assert ulps.max() <= 1
    finally:
        ctx.close()
