"""The autotuner's process-wide table of launch shapes: a saved table loads
back entry for entry, malformed or out-of-range lines are refused (host
only, no GPU).  The entries use compressions (123.25, 321.5) that no other
test ingests with, so loading them changes no other test's shape."""
import pytest


def _table(lines):
    return "# header\n" + "".join(l + "\n" for l in lines)


def test_table_round_trip(kd, tmp_path):
    src = tmp_path / "in.txt"
    lines = ["0 3 4638373364656766976 17 28 16 8192 32768 128 1",
             "0 2 4644363504004890624 10 20 8 4096 65536 64 0"]
    src.write_text(_table(lines))
    assert kd.GpuSession.autotune_load(str(src)) == 2
    out = tmp_path / "out.txt"
    kd.GpuSession.autotune_save(str(out))
    saved = [l for l in out.read_text().splitlines() if not l.startswith("#")]
    for l in lines:
        assert l in saved


@pytest.mark.parametrize("bad", ["0 3 1 2", "0 3 1 2 3 5 8192 32768 128 1",
                                 "0 9 1 2 3 16 8192 32768 128 1", "0 3 1 2 3 16 0 32768 128 1",
                                 "x y z"])
def test_table_rejects_bad_lines(kd, tmp_path, bad):
    src = tmp_path / "bad.txt"
    src.write_text(_table([bad]))
    with pytest.raises(kd.KdigestError):
        kd.GpuSession.autotune_load(str(src))


def test_table_missing_file(kd, tmp_path):
    with pytest.raises(kd.KdigestError):
        kd.GpuSession.autotune_load(str(tmp_path / "none.txt"))
