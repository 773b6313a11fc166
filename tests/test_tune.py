"""Autotune cache format and lookup, host only (cds_tuning_lookup reads the
file without touching a GPU)."""
import ctypes

import pytest

import cds_amd


def line(arch, cus, k, cls, shape=(256, 64), peel=0, shift=-1, replay=0, flow=0, ms=1.0):
    return (f"cds-tune 1 {arch} {cus} {k} {cls} {shape[0]} {shape[1]} {peel} {shift} "
            f"{replay} {flow} {ms:.4f}\n")


def test_struct_size():
    assert cds_amd.lib().cds_struct_size(4) == ctypes.sizeof(cds_amd.CTuning)


def test_nearest_class_within_two(tmp_path):
    p = tmp_path / "tune.txt"
    p.write_text("# comment\n\n"
                 + line("gfx950", 256, 3, 10, peel=8)
                 + line("gfx950", 256, 3, 20, shape=(512, 32), peel=128, shift=2)
                 + line("gfx950", 256, 4, 20, peel=64)
                 + line("gfx942", 304, 3, 20, peel=99)
                 + line("gfx950", 128, 3, 20, peel=77))
    look = lambda k, m, arch="gfx950", cus=256: cds_amd.tuning_lookup(str(p), arch, cus, k, m)
    # class = floor(log2 m): 2^20 .. 2^21-1 -> 20
    t = look(3, (1 << 20) + 5)
    assert t["peel_blocks"] == 128 and t["list_threads"] == 512 and t["list_chunk"] == 32
    assert t["peel_coarse_shift"] == 2
    assert look(3, 1 << 22)["peel_blocks"] == 128      # 2 classes away
    assert look(3, 1 << 23) is None                    # 3 classes away
    assert look(3, 1 << 11)["peel_blocks"] == 8        # class 11 -> entry 10
    assert look(3, 1 << 15) is None                    # 5 from both
    assert look(4, 1 << 19)["peel_blocks"] == 64       # per k
    assert look(2, 1 << 20) is None
    assert look(3, 1 << 20, arch="gfx942", cus=304)["peel_blocks"] == 99
    assert look(3, 1 << 20, cus=128)["peel_blocks"] == 77


def test_tie_goes_to_larger_class(tmp_path):
    p = tmp_path / "tune.txt"
    p.write_text(line("gfx950", 256, 3, 10, peel=1) + line("gfx950", 256, 3, 12, peel=2))
    assert cds_amd.tuning_lookup(str(p), "gfx950", 256, 3, 1 << 11)["peel_blocks"] == 2


@pytest.mark.parametrize("bad", [
    "cds-tune 2 gfx950 256 3 10 256 64 0 -1 0 0 1.0\n",      # version
    "cds-tune 1 gfx950 256 3 10 256 64 0 -1 0 0\n",          # field missing
    "cds-tune 1 gfx950 256 3 10 256 64 0 -1 0 0 1.0 7\n",    # field too many
    "cds-tune 1 gfx950 256 3 10 384 64 0 -1 0 0 1.0\n",      # no such shape
    "cds-tune 1 gfx950 256 3 10 128 64 0 -1 0 0 1.0\n",      # 128 lanes need 32
    "cds-tune 1 gfx950 256 3 10 256 64 0 9 0 0 1.0\n",       # coarse shift
    "cds-tune 1 gfx950 256 3 10 256 64 -2 -1 0 0 1.0\n",     # negative width
    "cds-tune 1 gfx950 256 9 10 256 64 0 -1 0 0 1.0\n",      # k
    "cds-tune 1 gfx950 256 3 10 256 64 0 -1 0 0 nan\n",      # time
    "tune 1 gfx950 256 3 10 256 64 0 -1 0 0 1.0\n",          # tag
])
def test_malformed_lines_are_refused(tmp_path, bad):
    p = tmp_path / "tune.txt"
    p.write_text(line("gfx950", 256, 3, 10) + bad)
    with pytest.raises(cds_amd.CdsError) as ei:
        cds_amd.tuning_lookup(str(p), "gfx950", 256, 3, 1 << 10)
    assert ei.value.status == 2 and "line 2" in str(ei.value)


def test_missing_file_is_an_io_error(tmp_path):
    with pytest.raises(cds_amd.CdsError) as ei:
        cds_amd.tuning_lookup(str(tmp_path / "none.txt"), "gfx950", 256, 3, 100)
    assert ei.value.status == 2


def test_tuning_dict_roundtrip():
    d = {"list_threads": 512, "list_chunk": 32, "peel_blocks": 64,
         "peel_coarse_shift": 0, "replay_blocks": 16, "flow_blocks": 8}
    assert cds_amd.CTuning.from_dict(d).as_dict() == d
    with pytest.raises(ValueError):
        cds_amd.CTuning.from_dict({"tile": 1})
