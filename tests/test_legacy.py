"""The legacy .dat converter (spinsim_amd/legacy.py), after the reference's
tools/dat2toml.py and docs/legacy.md:103-128: every archived @-script
converts and the drop-in's own `spinsim validate` accepts the result; the
CPMG and Rabi sequences agree step for step with the configurations the
reference's authors migrated by hand (configs/cpmg8z.toml,
configs/rabi_szsz.toml); synthetic scripts cover each keyword and error."""
import os
import pathlib
import subprocess

import pytest
import tomli

from spinsim_amd import legacy

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "spinsim_amd", "host", "build", "spinsim")
ARCHIVE = pathlib.Path("/root/reference/legacy/experiments")
CONFIGS = pathlib.Path("/root/reference/configs")


def scripts():
    """Archived @-scripts: .dat files whose every line is an @-directive."""
    out = []
    for p in sorted(ARCHIVE.rglob("*.dat")):
        lines = [ln.strip() for ln in p.read_text(errors="replace").splitlines() if ln.strip()]
        if lines and all(ln.startswith("@") for ln in lines):
            out.append(p)
    return out


def measurements(steps):
    n = 0
    for st in steps:
        if st["kind"] == "measure":
            n += 1
        elif st["kind"] == "repeat":
            n += st["count"] * measurements(st["body"])
    return n


def effective(step, base):
    """(tau x steps, hx, hy, hz) of an evolve step, fields falling back to the baseline."""
    return (step["tau"] * step.get("steps", 1),) + tuple(step.get(k, base[k]) for k in ("hx", "hy", "hz"))


def validate(tmp_path, text, name="c.toml"):
    if not os.path.exists(CLI):
        pytest.fail(f"{CLI} missing: run __graft_entry__.build()")
    p = tmp_path / name
    p.write_text(text)
    return subprocess.run([CLI, "validate", str(p)], capture_output=True, text=True, timeout=60)


def test_archive_converts_and_validates(tmp_path):
    if not ARCHIVE.is_dir():
        pytest.skip("the reference's legacy archive is not present")
    found = scripts()
    assert len(found) >= 100, len(found)
    for path in found:
        text = legacy.convert(path)
        doc = tomli.loads(text)
        r = validate(tmp_path, text)
        assert r.returncode == 0, (path, r.stderr)
        # one measurement up front, one per cycle
        repeats = [s for s in doc["step"] if s["kind"] == "repeat"]
        want = 1 + (repeats[0]["count"] if repeats else 0)
        assert measurements(doc["step"]) == want, path


def _body(doc):
    rep = [s for s in doc["step"] if s["kind"] == "repeat"]
    assert len(rep) == 1
    return rep[0]


@pytest.mark.parametrize("legacy_dir,config", [
    ("cpmg/CPMGZ/CPMG8Z/nmrCPMG8.dat", "cpmg8z.toml"),
    ("rabi/Rabi_10_SzSz/nmrRabi.dat", "rabi_szsz.toml"),
])
def test_sequences_match_hand_migrated_configs(legacy_dir, config):
    """The evolve steps of the cycle (duration and the fields in force)
    agree with the reference's hand-written configuration of the same
    experiment; counts, spins and ensemble settings were edited by hand there
    and are not compared."""
    src, ref = ARCHIVE / legacy_dir, CONFIGS / config
    if not (src.exists() and ref.exists()):
        pytest.skip("the reference's legacy archive or configs are not present")
    got = tomli.loads(legacy.convert(src))
    want = tomli.loads(ref.read_text())
    gb, wb = got["field"], want["field"]
    g = [effective(s, gb) for s in _body(got)["body"] if s["kind"] == "evolve"]
    w = [effective(s, wb) for s in _body(want)["body"] if s["kind"] == "evolve"]
    assert g == w
    assert [s["kind"] for s in _body(got)["body"]][-1] == "measure"


def write_step(d, name, spins, tau, nsteps, intermediate=1, initial=None, extra=""):
    rows = "".join(f"SPIN : {k} (Hx0, Hx1, Omega_x, Hy0, Hy1, Omega_y, Hz0, Hz1, Omega_z, phi_x, phi_y, phi_z)\n"
                   f"{hx}, 0, 0, {hy}, 0, 0, {hz}, 0, 0, 0, 0, 0,\nEND\n" for k, (hx, hy, hz) in spins.items())
    text = ("HEISENBERG INTERACTION\n1, 4, 0.0, 0.0, 2.5,\nEND\n" + rows +
            f"NUMBER OF FIELD STEPS\n{nsteps}\nEND\nTIME STEP ;\n{tau}\nEND\n"
            f"NUMBER OF INTERMEDIATE TIME STEPS\n{intermediate}\nEND\n")
    if initial:
        text += f"INITIAL STATE: {initial}\n"
    (d / name).write_text(text + extra)


def test_synthetic_script(tmp_path):
    write_step(tmp_path, "setup.dat", {1: (0.5, 0.0, 1.0)}, 0.01, 0, initial="SPIN DOWN")
    write_step(tmp_path, "free.dat", {1: (0.0, 0.0, 0.0)}, 0.02, 3, intermediate=5)
    write_step(tmp_path, "kick.dat", {k: (10.0 * k, 0.0, -1.0) for k in range(1, 5)}, 0.001, 1)
    write_step(tmp_path, "cool.dat", {1: (0.0, 0.0, 0.0)}, 0.1, 2, extra="IMAGINARY TIME PROPAGATION\n")
    (tmp_path / "run.dat").write_text("@./setup.dat\n@CYCLE 7\n@./free.dat\n@./kick.dat\n@ENDCYCLE\n@./cool.dat\n")
    text = legacy.convert(tmp_path / "run.dat")
    doc = tomli.loads(text)
    assert doc["system"] == {"spins": 4, "initial_state": "down"}
    assert doc["field"] == {"hx": 0.5, "hy": 0.0, "hz": 1.0}
    assert doc["ensemble"]["dipolar_scale"] == 2.5
    steps = doc["step"]
    assert [s["kind"] for s in steps] == ["measure", "repeat", "evolve"]
    rep = steps[1]
    assert rep["count"] == 7
    free, kick, meas = rep["body"]
    # NUMBER OF INTERMEDIATE TIME STEPS multiplies tau; zero fields are written out
    assert free == {"kind": "evolve", "tau": 0.1, "steps": 3, "hx": 0.0, "hy": 0.0, "hz": 0.0}
    assert kick["hx"] == [10.0, 20.0, 30.0, 40.0] and kick["hz"] == -1.0
    assert meas == {"kind": "measure"}
    assert steps[2]["mode"] == "imaginary" and steps[2]["steps"] == 2
    assert "WARNING: cool.dat: imaginary-time step" in text
    assert validate(tmp_path, text).returncode == 0


def test_partial_spin_rows_fall_back_with_a_warning(tmp_path):
    write_step(tmp_path, "a.dat", {1: (1.0, 0.0, 0.0), 2: (2.0, 0.0, 0.0)}, 0.1, 1)   # 4 spins, 2 rows
    (tmp_path / "s.dat").write_text("@./a.dat\n")
    text = legacy.convert(tmp_path / "s.dat")
    doc = tomli.loads(text)
    assert doc["step"][1]["hx"] == 1.0
    assert "fields differ between spins" in text


@pytest.mark.parametrize("script,message", [
    ("@CYCLE 2\n@./a.dat\n@CYCLE 3\n@./a.dat\n@ENDCYCLE\n@ENDCYCLE\n", "second @CYCLE"),
    ("@./a.dat\n@ENDCYCLE\n", "@ENDCYCLE without @CYCLE"),
    ("@./missing.dat\n", "does not exist"),
    ("\n", "includes no step files"),
])
def test_script_errors(tmp_path, script, message):
    write_step(tmp_path, "a.dat", {1: (0.0, 0.0, 0.0)}, 0.1, 1)
    (tmp_path / "s.dat").write_text(script)
    with pytest.raises(legacy.LegacyError, match=message):
        legacy.convert(tmp_path / "s.dat")
    assert legacy.main([str(tmp_path / "s.dat")]) == 1


def test_short_spin_row(tmp_path):
    (tmp_path / "a.dat").write_text("SPIN : 1 (Hx0, ...)\n1.0, 0, 0\nEND\n")
    (tmp_path / "s.dat").write_text("@./a.dat\n")
    with pytest.raises(legacy.LegacyError, match="has 3 of 9 columns"):
        legacy.convert(tmp_path / "s.dat")
