"""The host driver (lbm-c_amd/lbmc_run, the index.cu main loop on the C ABI)
end to end: .lbmc + field file in, solutions out; FP64 storage matches the
oracle bit for bit; the steady-state stop, the NaN abort and checkpoint
restart behave as the reference's loop (LBM-C/index.cu:102-142)."""
import os
import subprocess

import numpy as np
import pytest

import cases
import fieldfile

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "lbm-c_amd", "lbmc_run")
import pyoracle  # noqa: E402


def write_case(tmp, Lx=32, Ly=16, maxt=300, fileout=100, screen=50, steady=0, tol=0.0, init=0,
               domain="in.lbmf", out="out.lbmf", force=1e-5, nan_at=None, storage="FP64"):
    geo = np.zeros((Ly, Lx))
    geo[0, :] = 1
    geo[-1, :] = 1
    fx = np.full(Lx * Ly, force)
    if nan_at is not None:
        fx[nan_at] = np.nan
    fieldfile.write(tmp / domain, {"Porosity": geo.reshape(-1), "ForceX": fx, "ForceY": np.zeros(Lx * Ly)})
    (tmp / "case.lbmc").write_text(f"""ProjName chan
Lattice D2Q9
Storage {storage}
DomainFile {tmp / domain}
OutputFile {tmp / out}
Tau 0.8
Lx {Lx}
Ly {Ly}
ColType BGK
Force 1
MicroBC 0
MacroBC 0
Tolerance {tol}
Init {init}
MaxT {maxt}
FileOut {fileout}
ScreenMes {screen}
SteadyCheck {steady}
OutputVars 1 1 0 1 0
ScreenNode 5 8 0
Interactive 0
""")
    return tmp / "case.lbmc"


def run(case, *args):
    print(f"[driver] {case} {args}", flush=True)
    return subprocess.run([RUN, str(case), *args], capture_output=True, text=True, timeout=120)


def oracle_channel(Lx, Ly, steps, force=1e-5):
    geo = np.zeros((Ly, Lx))
    geo[0, :] = 1
    geo[-1, :] = 1
    o = pyoracle.Oracle(pyoracle.D2Q9, (Lx, Ly), pyoracle.BGK, forcing=True, tau=0.8)
    o.set_fields(geometry=geo.reshape(-1), force=[np.full(Lx * Ly, force), np.zeros(Lx * Ly)])
    o.init_static()
    o.run(0, steps, steps)
    return o


def test_driver_matches_oracle(tmp_path):
    case = write_case(tmp_path)
    r = run(case)
    assert r.returncode == 0, r.stderr
    assert "time = 300; rho = " in r.stdout
    out = fieldfile.read(tmp_path / "out.lbmf")
    assert list(out["Dims"]) == [32, 16, 1]
    assert {"VelocityX@100", "VelocityY@200", "Density@300"} <= set(out)
    o = oracle_channel(32, 16, 300)
    rho, u = o.get_macros()
    assert np.array_equal(out["VelocityX@300"], u[0])
    assert np.array_equal(out["VelocityY@300"], u[1])
    assert np.array_equal(out["Density@300"], rho)


def test_driver_steady_stop(tmp_path):
    # three residuals below a large tolerance: stops at the first check
    case = write_case(tmp_path, maxt=1000, fileout=0, screen=0, steady=40, tol=10.0)
    r = run(case)
    assert r.returncode == 0, r.stderr
    assert "Steady state reached at time 40" in r.stdout
    assert "VelocityX@40" in fieldfile.read(tmp_path / "out.lbmf")


def test_driver_nan_abort(tmp_path):
    case = write_case(tmp_path, maxt=1000, fileout=0, screen=0, steady=20, tol=0.0, nan_at=100)
    r = run(case)
    assert r.returncode == 1, r.stdout + r.stderr
    assert "Residual is NaN at time 20" in r.stdout
    assert "VelocityX@20" in fieldfile.read(tmp_path / "out.lbmf")


def test_driver_missing_field(tmp_path):
    case = write_case(tmp_path)
    fieldfile.write(tmp_path / "in.lbmf", {"Porosity": np.zeros(32 * 16)})
    r = run(case)
    assert r.returncode != 0
    assert 'Input Handler: ForceX not found in file' in r.stdout


def test_driver_checkpoint_restart(tmp_path):
    full = write_case(tmp_path, maxt=300, fileout=300, screen=0, out="full.lbmf")
    assert run(full).returncode == 0
    first = write_case(tmp_path, maxt=180, fileout=0, screen=0, out="a.lbmf")
    r = run(first, "--checkpoint", str(tmp_path / "ck.lbmf"))
    assert r.returncode == 0, r.stderr
    ck = fieldfile.read(tmp_path / "ck.lbmf")
    assert ck["TimeStep"][0] == 180 and "Porosity" in ck and "f8" in ck
    second = write_case(tmp_path, maxt=120, fileout=120, screen=0, init=2, domain="ck.lbmf", out="b.lbmf")
    # write_case rewrote ck.lbmf's inputs: restore the checkpoint records
    fieldfile.write(tmp_path / "ck.lbmf", ck)
    r = run(second)
    assert r.returncode == 0, r.stderr
    a = fieldfile.read(tmp_path / "full.lbmf")["VelocityX@300"]
    b = fieldfile.read(tmp_path / "b.lbmf")["VelocityX@120"]
    assert np.array_equal(a, b)


def test_driver_cgns_in_and_out(tmp_path):
    """DomainFile and OutputFile as CGNS (through libcgnsio.so): same result
    as the field-file run and the oracle."""
    import test_cgns_io as cg

    if not (os.path.exists(cg.LIB) and os.path.exists(cg.H5DUMP)):
        pytest.skip("libcgnsio.so or h5dump not available")
    Lx, Ly = 32, 16
    geo = np.zeros((Ly, Lx))
    geo[0, :] = 1
    geo[-1, :] = 1
    L = cg.lib()
    fin = str(tmp_path / "in.cgns")
    assert L.cgio_write_grid(fin.encode(), Lx, Ly, 1) == 0
    cg.append(L, fin, 0, {"Porosity": geo.reshape(-1), "ForceX": np.full(Lx * Ly, 1e-5),
                          "ForceY": np.zeros(Lx * Ly)})
    case = write_case(tmp_path, maxt=300, fileout=150, screen=0, domain="unused.lbmf", out="out.cgns")
    text = case.read_text().replace(str(tmp_path / "unused.lbmf"), fin)
    case.write_text(text)
    r = run(case)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Input Handler: Porosity loaded" in r.stdout
    fout = str(tmp_path / "out.cgns")
    shape, ux = cg.dump_data(fout, "/Base/LBM-C Output Data/Solution @ Time = 300/VelocityX/ data")
    assert shape == (Ly, Lx)
    o = oracle_channel(Lx, Ly, 300)
    _, u = o.get_macros()
    # h5dump prints enough digits to tell FP64 values apart at this size
    assert np.allclose(ux, u[0], rtol=1e-12, atol=1e-18)
    assert list(cg.dump_data(fout, "/Base/TimeIterValues/IterationValues/ data")[1]) == [150, 300]
