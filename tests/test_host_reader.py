"""The .lbmc reader (LBM-C/infile_reader.cu:33-248): keys matched by
substring in the reference's order, '#' comments, stream extraction of the
values, unknown lines echoed."""
import json
import os
import subprocess

import numpy as np
import pytest

import fieldfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PARSE = os.path.join(ROOT, "lbm-c_amd", "lbmc_parse")

pytestmark = pytest.mark.skipif(not os.path.exists(PARSE), reason="lbmc_parse not built")


def parse(tmp_path, text):
    p = tmp_path / "case.lbmc"
    p.write_text(text)
    r = subprocess.run([PARSE, str(p)], capture_output=True, text=True, check=True)
    return json.loads(r.stdout), r.stderr


def test_full_key_set(tmp_path):
    cfg, echo = parse(tmp_path, """# cylinder case
ProjName   cyl
DomainFile cyl_in.lbmf
OutputFile cyl_out.lbmf
TauMRT 0 1.64 1.54 0 1.9 0 1.9 1.25 1.25 0 0 0 0 0 0
Tau 0.8   # relaxation time
Lx 256
Ly 64
Lz 1
DeltaX 1
DeltaT 1
C_smag 0.17
ColType MRTPOR
Force 1
MicroBC 1
MacroBC 0
Tolerance 1e-9
Init 0
MaxT 10000
FileOut 1000
ScreenMes 500
SteadyCheck 100
OutputVars 1 1 0 1 0
ScreenNode 128 32 0
Interactive 0
something unknown
""")
    assert cfg["ProjName"] == "cyl" and cfg["DomainFile"] == "cyl_in.lbmf"
    assert cfg["OutputFile"] == "cyl_out.lbmf"
    assert cfg["tau"] == 0.8
    assert cfg["tau_mrt"][:9] == [0, 1.64, 1.54, 0, 1.9, 0, 1.9, 1.25, 1.25]
    assert cfg["length"] == [256, 64, 1]  # D3Q15 default reads Lz
    assert cfg["c_smag"] == 0.17
    assert cfg["collision_type"] == 3      # MRTPOR overrides MRT
    assert cfg["forcing"] == 1 and cfg["micro_bc"] == 1 and cfg["macro_bc"] == 0
    assert cfg["tolerance"] == 1e-9 and cfg["init_type"] == 0
    assert (cfg["max"], cfg["plot"], cfg["screen"], cfg["steady_check"]) == (10000, 1000, 500, 100)
    assert cfg["out_u"] == [1, 1, 0] and cfg["out_rho"] == 1 and cfg["out_pressure"] == 0
    assert cfg["screen_node"] == [128, 32, 0]
    assert "something unknown" in echo
    assert "ColType = MRTPOR" in echo


def test_substring_order_and_comments(tmp_path):
    cfg, _ = parse(tmp_path, """Lattice D2Q9
TauMRT 0 1 2 3 4 5 6 7 8
Tau 0.6 # TauMRT on a comment does not matter
# Tau 5.0   whole-line comment
Lz 7
ColType BGK
""")
    assert cfg["lattice"] == 0
    assert cfg["tau"] == 0.6
    assert cfg["tau_mrt"][:9] == list(range(9)) and cfg["tau_mrt"][9] == 0
    assert cfg["length"][2] == 0           # Lz is ignored for D2Q9
    assert cfg["collision_type"] == 0


@pytest.mark.parametrize("name,code", [("BGK", 0), ("NTPOR", 1), ("MRT", 2), ("MRTPOR", 3)])
def test_coltype_codes(tmp_path, name, code):
    cfg, _ = parse(tmp_path, f"ColType {name}\n")
    assert cfg["collision_type"] == code


def test_storage_key(tmp_path):
    for v, code in (("FP64", 0), ("FP32", 1), ("FP16S", 2), ("BF16S", 3)):
        cfg, _ = parse(tmp_path, f"Storage {v}\n")
        assert cfg["storage"] == code


def test_fieldfile_roundtrip(tmp_path):
    p = tmp_path / "f.lbmf"
    rng = np.random.default_rng(0)
    fields = {"Porosity": rng.random(12), "MicroBC": np.arange(12, dtype=np.int32)}
    fieldfile.write(p, fields)
    back = fieldfile.read(p)
    assert np.array_equal(back["Porosity"], fields["Porosity"])
    assert np.array_equal(back["MicroBC"], fields["MicroBC"]) and back["MicroBC"].dtype == np.int32
