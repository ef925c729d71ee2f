"""Runs the C++ drop-in's test driver (tests/cpp/test_host_api.cpp) and checks
the C++ realisation recipe against the oracle, realisation by realisation."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "spinsim_amd", "host", "build", "test_host_api")


def driver(*args, timeout=240):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} missing: run __graft_entry__.build()")
    print(f"[host] {' '.join(args)}", flush=True)
    r = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-4000:], r.stderr[-2000:])
    return r


def test_host_api_cpu_cases():
    r = driver("cpu")
    assert r.returncode == 0 and "OK: 0 failure(s)" in r.stdout


REFERENCE_CONFIGS = "/root/reference/configs"


def measurements(steps):
    n = 0
    for st in steps:
        if st["kind"] == "measure":
            n += 1
        elif st["kind"] == "repeat":
            n += st["count"] * measurements(st["body"])
    return n


@pytest.mark.parametrize("name", ["cpmg8z.toml", "rabi_szsz.toml"])
def test_reference_configs_load_unchanged(name):
    """The reference's own configuration files through the drop-in's
    loadConfig; the expected shape comes from an independent TOML reader."""
    path = os.path.join(REFERENCE_CONFIGS, name)
    if not os.path.exists(path):
        pytest.skip("the reference's configs are not present")
    import tomli
    with open(path, "rb") as f:
        doc = tomli.load(f)
    want = (f"spins={doc['system']['spins']} realizations={doc['ensemble']['realizations']} "
            f"seed={doc['ensemble']['seed']} observables={len(doc['measurement']['observables'])} "
            f"measurements={measurements(doc['step'])} steps={len(doc['step'])}")
    r = driver("config", path)
    assert r.returncode == 0 and r.stdout.strip() == want, (r.stdout, want)


@pytest.mark.gpu
def test_host_api_device_cases():
    r = driver("gpu")
    assert r.returncode == 0 and "OK: 0 failure(s)" in r.stdout


@pytest.mark.gpu
def test_cpp_recipe_matches_oracle_per_realisation(tmp_path):
    """Config-0 CPMG echo (L=10, 20 realisations of seed 20240726) through the
    C++ API: Rng, geometry (min spacing 0.5, (1-3cos^2)/r^3), disorder and
    initial state drawn on the host, propagation on the device. Each
    realisation's curve is within 1e-10 of the oracle's, and the batched
    runner's ensemble mean agrees with the oracle's fold."""
    out = tmp_path / "echo.txt"
    r = driver("dump", str(out))
    assert r.returncode == 0
    rows, means = [], []
    for line in out.read_text().splitlines():
        parts = line.split()
        if parts[0] == "mean":
            means.append(float(parts[1]))
        else:
            rows.append([float(v) for v in parts])
    rows = np.array(rows)
    L, n = 10, 20
    p = O.Program()
    allspins = list(range(L))
    p.pulse((1, 0, 0), np.pi / 2, allspins).repeat(8).evolve(0.05, 10)
    p.pulse((0, 1, 0), np.pi, allspins).evolve(0.05, 10).measure().end()
    cfg = O.Config(L=L, program=p, observables=((O.OBS_CORR_X, 0), (O.OBS_NORM, 0), (O.OBS_ENERGY, 0)),
                   field_disorder=0.5, min_separation=0.5, angular=True)
    times, per, mean, _, _ = O.run_ensemble(cfg, n, threads=8)
    got = rows[:, 2:].reshape(n, 8, 5)
    assert np.array_equal(rows[:8, 1], times)
    assert float(np.max(np.abs(got - per))) < 1e-10
    assert float(np.max(np.abs(np.array(means).reshape(8, 5) - mean))) < 1e-10
