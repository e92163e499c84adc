"""The autotuner's table (CPU only: nnet_tuning_read parses on the host).

Settings override the defaults one by one; every malformed line is rejected with NNET_E_ARG and
leaves nothing applied; the committed table, which nnet_ctx_create loads by default, parses."""
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "sph-exa-nnet_amd", "data", "tuning_gfx950.txt")


def _nm():
    import nnet_mi355x
    if not os.path.exists(nnet_mi355x.LIB_PATH):
        pytest.skip("libnnet_mi355x.so not built")
    return nnet_mi355x


def test_defaults_and_overrides(tmp_path):
    nm = _nm()
    p = tmp_path / "t.txt"
    p.write_text("# comment line\n\nppw 14 1 0 4   # net14, Helmholtz, FP64\nppw 14 0 1 2\nppw 86 1 1 1\nfused 0\n")
    t = nm.tuning_read(str(p))
    assert t["ppw"][(14, 1, 0)] == 4 and t["ppw"][(14, 0, 1)] == 2
    assert t["ppw"][(14, 0, 0)] == 1 and t["ppw"][(14, 1, 1)] == 1  # not listed: default
    assert all(t["ppw"][(k, e, q)] == 1 for k in (86, 87) for e in (0, 1) for q in (0, 1))
    assert t["fused"] is False
    p.write_text("")
    t = nm.tuning_read(str(p))
    assert t["fused"] is True and set(t["ppw"].values()) == {1}


@pytest.mark.parametrize("line", ["ppw 86 1 0 2", "ppw 14 1 0 3", "ppw 14 2 0 1", "ppw 14 1 2 1", "ppw 15 1 0 1",
                                  "ppw 14 1 0", "ppw 14 1 0 2 7", "fused 2", "fused", "lu_block 8", "ppw x 1 0 1"])
def test_malformed_lines_rejected(tmp_path, line):
    nm = _nm()
    p = tmp_path / "t.txt"
    p.write_text("ppw 14 1 0 4\n" + line + "\n")
    with pytest.raises(nm.NnetError, match="-1"):
        nm.tuning_read(str(p))


def test_missing_table_rejected(tmp_path):
    nm = _nm()
    with pytest.raises(nm.NnetError, match="-1"):
        nm.tuning_read(str(tmp_path / "absent.txt"))


def test_committed_table_parses():
    nm = _nm()
    if not os.path.exists(TABLE):
        pytest.skip("no committed tuning table")
    t = nm.tuning_read(TABLE)
    assert set(t["ppw"][(14, e, q)] for e in (0, 1) for q in (0, 1)) <= {1, 2, 4}
