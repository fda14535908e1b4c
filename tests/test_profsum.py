"""The rocprof summary tool's derived counter table (nbhip.profsum
--counters) on small synthetic summaries: issue use and resident waves follow
from the counters and the average launch time as the table's header states."""
from tests import helpers  # noqa: F401  (puts the package on sys.path)
from nbhip import profsum

PASS1 = """# counters pass 1

per dispatch (sum / dispatches):

| kernel | SQ_INSTS_LDS | SQ_INSTS_SALU | SQ_INSTS_VALU | SQ_INSTS_VALU_ADD_F64 | SQ_INSTS_VALU_FMA_F64 | SQ_INSTS_VALU_MUL_F64 |
|---|---|---|---|---|---|---|
| nbh::k_mid | 100 | 200 | 614400 | 153600 | 76800 | 76800 |
"""
PASS2 = """# counters pass 2

per dispatch (sum / dispatches):

| kernel | SQ_WAVE_CYCLES |
|---|---|
| nbh::k_mid | 1536000 |
"""
TRACE = """| kernel | calls | total us | avg us | min us | max us | % |
|---|---|---|---|---|---|---|
| nbh::k_mid | 2 | 20.0 | 10.00 | 9.0 | 11.0 | 100.0 |
"""


def test_counters_report(tmp_path):
    p1, p2, tr = tmp_path / "p1.md", tmp_path / "p2.md", tmp_path / "t.md"
    p1.write_text(PASS1)
    p2.write_text(PASS2)
    tr.write_text(TRACE)
    out = profsum.counters_report(str(p1), str(p2), str(tr), ["cmd1", "cmd2"], ["nbh::k_mid"])
    row = [x for x in out.split("\n") if x.startswith("| k_mid |")][0]
    cells = [c.strip() for c in row.strip("|").split("|")]
    # 10 us at 2.4 GHz = 24000 cycles; 256 CUs x 4 SIMDs x 24000 / 4 = 6.144e6 issue slots
    assert cells[1] == "10.0"
    assert cells[3] == "50%"             # FP64 share: 307200 / 614400
    assert cells[4] == "10%"             # VALU issue: 614400 / 6.144e6
    assert cells[5] == "5%"              # FP64 issue
    assert cells[8] == "1.0"             # waves/CU: 1536000 * 4 / 24000 / 256
    assert "cmd1" in out and "cmd2" in out
