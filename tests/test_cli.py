"""The spinsim command line of the C++ drop-in (spinsim_amd/host/apps/spinsim.cpp),
after the reference's apps/spinsim/main.cpp:17-158: commands, options, exit
codes and messages on the CPU; `run` on the device against the oracle's
ensemble and, for fixed coupling tables, against the dense
eigendecomposition reference."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
import dense as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "spinsim_amd", "host", "build", "spinsim")
REFERENCE_CONFIGS = "/root/reference/configs"

CPMG = """
[system]
spins = {L}
[ensemble]
realizations = {n}
seed = 20240726
field_disorder = 0.5
[measurement]
reference_axis = "x"
observables = ["corr:x", "norm", "energy"]
[[step]]
kind = "pulse"
rotation = "pi/2:x"
[[step]]
kind = "repeat"
count = 4
body = [
  {{ kind = "evolve", tau = 0.05, steps = 10 }},
  {{ kind = "pulse", rotation = "pi:y" }},
  {{ kind = "evolve", tau = 0.05, steps = 10 }},
  {{ kind = "measure" }},
]
"""

# an XXZ chain with one anisotropic long-range pair, fields along x and z,
# started from a basis state: no random draws at all
CHAIN_L = 6
CHAIN_HZ = [0.2, -0.1, 0.3, 0.0, 0.5, -0.4]
CHAIN_HX = 0.15
CHAIN_PAIRS = [((1, 2), 1.0, 1.0, 0.6), ((2, 3), 1.0, 1.0, 0.6), ((3, 4), 1.0, 1.0, 0.6),
               ((4, 5), 1.0, 1.0, 0.6), ((5, 6), 1.0, 1.0, 0.6), ((1, 4), 0.3, -0.2, 0.0)]
CHAIN_BASIS = 0b010101


def chain_config(realizations=2, geometry=False, extra_pairs=()):
    lines = [
        "[system]", f"spins = {CHAIN_L}", 'initial_state = "basis"', f"basis_index = {CHAIN_BASIS}",
        "[field]", f"hx = {CHAIN_HX}", f"hz = {CHAIN_HZ}",
        "[ensemble]", f"realizations = {realizations}", f"dipolar_geometry = {'true' if geometry else 'false'}",
        "[measurement]", 'observables = ["mag:1:z", "mag:4:z", "energy"]',
        "[numerics]", "epsilon = 1e-12",
    ]
    for (i, j), jx, jy, jz in list(CHAIN_PAIRS) + list(extra_pairs):
        lines += ["[[coupling]]", f"spins = [{i}, {j}]", f"jx = {jx}", f"jy = {jy}", f"jz = {jz}"]
    lines += ["[[step]]", 'kind = "repeat"', "count = 10",
              'body = [{ kind = "evolve", tau = 0.1, steps = 1 }, { kind = "measure" }]']
    return "\n".join(lines) + "\n"


def spinsim(*args, timeout=300):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} missing: run __graft_entry__.build()")
    print(f"[cli] spinsim {' '.join(map(str, args))}", flush=True)
    r = subprocess.run([BIN, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-3000:], r.stderr[-2000:])
    return r


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


def read_csv(text):
    lines = text.strip().splitlines()
    header = lines[0].split(",")
    rows = np.array([[float(v) for v in line.split(",")] for line in lines[1:]])
    return header, rows


# ---------------------------------------------------------------- CPU cases

def test_usage_and_exit_codes():
    r = spinsim()
    assert r.returncode == 2 and "Usage:" in r.stdout
    for flag in ("-h", "--help"):
        r = spinsim(flag)
        assert r.returncode == 0 and "spinsim run <config.toml>" in r.stdout


@pytest.mark.parametrize("args,message", [
    (["run"], "no command given"),
    (["frob", "{cfg}"], "unknown command 'frob'"),
    (["run", "{cfg}", "--threads"], "--threads needs a value"),
    (["run", "{cfg}", "--threads", "two"], "--threads expects a non-negative integer, got 'two'"),
    (["run", "{cfg}", "--realizations", "-1"], "--realizations expects a non-negative integer"),
    (["run", "{cfg}", "--fast"], "unknown option '--fast'"),
    (["run", "{cfg}", "--backend", "cpu"], "the only backend is hip"),
    (["validate", "{cfg}", "--realizations", "0"], "ensemble.realizations"),
    (["validate", "{missing}"], "cannot be opened"),
])
def test_argument_errors(tmp_path, args, message):
    cfg = write(tmp_path, "c.toml", CPMG.format(L=6, n=3))
    args = [a.format(cfg=cfg, missing=tmp_path / "none.toml") for a in args]
    r = spinsim(*args)
    assert r.returncode == 1 and r.stderr.startswith("spinsim: ") and message in r.stderr


def test_validate_and_describe(tmp_path):
    cfg = write(tmp_path, "c.toml", CPMG.format(L=6, n=3))
    r = spinsim("validate", cfg)
    assert r.returncode == 0 and r.stdout == f"{cfg}: ok\n"
    r = spinsim("describe", cfg, "--realizations", "7")
    assert r.returncode == 0
    want = ["spins           6", "state space     64 amplitudes", "initial state   random",
            "realizations    7", "seed            20240726", "geometry        dipolar 2D, per realization",
            "field disorder  0.5", "reference axis  x", "epsilon         1e-07",
            "measurements    4 per realization", "observables     corr:x norm energy"]
    assert r.stdout.splitlines() == want


@pytest.mark.parametrize("name", ["cpmg8z.toml", "rabi_szsz.toml"])
def test_reference_configs_validate_and_describe(name):
    path = os.path.join(REFERENCE_CONFIGS, name)
    if not os.path.exists(path):
        pytest.skip("the reference's configs are not present")
    import tomli
    with open(path, "rb") as f:
        doc = tomli.load(f)
    assert spinsim("validate", path).stdout == f"{path}: ok\n"
    r = spinsim("describe", path)
    got = dict(line.split(None, 1) if not line.startswith(("state", "initial", "field", "reference"))
               else (" ".join(line.split()[:2]), line.split(None, 2)[2]) for line in r.stdout.splitlines())
    assert got["spins"] == str(doc["system"]["spins"])
    assert got["realizations"] == str(doc["ensemble"]["realizations"])
    assert got["seed"] == str(doc["ensemble"]["seed"])
    assert got["observables"].split() == doc["measurement"]["observables"]


def test_coupling_tables(tmp_path):
    cfg = write(tmp_path, "chain.toml", chain_config())
    r = spinsim("describe", cfg)
    assert r.returncode == 0 and "geometry        from config, 6 fixed couplings" in r.stdout
    for text, message in [
        (chain_config(geometry=True), "fixed couplings need ensemble.dipolar_geometry = false"),
        (chain_config(extra_pairs=[((2, 1), 0.1, 0.1, 0.0)]), "pair 7 repeats pair 1"),
        (chain_config(extra_pairs=[((1, 7), 0.1, 0.1, 0.0)]), "between 1 and 6"),
        (chain_config(extra_pairs=[((3, 3), 0.1, 0.1, 0.0)]), "between 1 and 6"),
        (chain_config(extra_pairs=[((0, 2), 0.1, 0.1, 0.0)]), "between 1 and 6"),
    ]:
        bad = write(tmp_path, "bad.toml", text)
        r = spinsim("validate", bad)
        assert r.returncode == 1 and "'coupling" in r.stderr and message in r.stderr, r.stderr


# ------------------------------------------------------------- device cases

def oracle_cpmg(L):
    p = O.Program()
    allspins = list(range(L))
    p.pulse((1, 0, 0), np.pi / 2, allspins).repeat(4).evolve(0.05, 10)
    p.pulse((0, 1, 0), np.pi, allspins).evolve(0.05, 10).measure().end()
    cfg = O.Config(L=L, program=p, observables=((O.OBS_CORR_X, 0), (O.OBS_NORM, 0), (O.OBS_ENERGY, 0)),
                   field_disorder=0.5, reference_axis=0)   # 0 = x
    return cfg


@pytest.mark.gpu
@pytest.mark.parametrize("batched", [False, True])
def test_run_matches_the_oracle_ensemble(tmp_path, batched):
    """`spinsim run` on an L = 8 CPMG echo of 6 realisations: the CSV's mean,
    sd and stderr columns agree with the oracle's ensemble to the CSV's 12
    significant digits (1e-10 between the device and the oracle)."""
    L, n = 8, 6
    cfg = write(tmp_path, "cpmg.toml", CPMG.format(L=L, n=n))
    out = tmp_path / "out.csv"
    args = ["run", cfg, "--threads", "2", "--out", out] + (["--batched"] if batched else [])
    r = spinsim(*args)
    assert r.returncode == 0 and f"{n} realizations on 2 threads" in r.stderr
    header, rows = read_csv(out.read_text())
    assert header[:4] == ["time", "corr_x_re_mean", "corr_x_re_sd", "corr_x_re_stderr"]
    times, _, mean, sd, se = O.run_ensemble(oracle_cpmg(L), n, threads=4)
    assert np.allclose(rows[:, 0], times, rtol=1e-9, atol=0)
    width = mean.shape[1]
    for k, want in enumerate((mean, sd, se)):
        got = rows[:, 1 + k::3]
        assert got.shape == want.shape == (4, width)
        assert np.allclose(got, want, rtol=1e-9, atol=1e-10), (k, np.max(np.abs(got - want)))


@pytest.mark.gpu
def test_run_to_stdout_equals_out_file(tmp_path):
    cfg = write(tmp_path, "cpmg.toml", CPMG.format(L=6, n=3))
    out = tmp_path / "out.csv"
    a = spinsim("run", cfg, "--threads", "1")
    b = spinsim("run", cfg, "--threads", "3", "--out", out)
    assert a.returncode == 0 and b.returncode == 0
    assert a.stdout == out.read_text()


@pytest.mark.gpu
def test_coupling_tables_match_the_dense_reference(tmp_path):
    """A fixed XXZ chain from [[coupling]] tables (L = 6, basis initial
    state): mag:1:z, mag:4:z and the energy at t = 0.1 .. 1.0 agree with
    exp(iHt) from the dense Hermitian eigendecomposition."""
    cfg = write(tmp_path, "chain.toml", chain_config())
    r = spinsim("run", cfg)
    assert r.returncode == 0
    header, rows = read_csv(r.stdout)
    means = rows[:, [header.index(c) for c in ("mag_1_z_mean", "mag_4_z_mean", "energy_mean")]]
    fields = np.zeros((CHAIN_L, 3))
    fields[:, 0] = CHAIN_HX
    fields[:, 2] = CHAIN_HZ
    lo = [i - 1 for (i, _), *_ in CHAIN_PAIRS]
    hi = [j - 1 for (_, j), *_ in CHAIN_PAIRS]
    H = D.hamiltonian(CHAIN_L, fields, lo, hi, [p[1] for p in CHAIN_PAIRS], [p[2] for p in CHAIN_PAIRS],
                      [p[3] for p in CHAIN_PAIRS])
    psi0 = np.zeros(2 << CHAIN_L)
    psi0[2 * CHAIN_BASIS] = 1.0
    z0, z3 = D.pauli(CHAIN_L, 0, "z"), D.pauli(CHAIN_L, 3, "z")
    for k in range(10):
        t = 0.1 * (k + 1)
        c = D.to_complex(D.expm_apply(H, 1j * t, psi0))
        want = [np.vdot(c, z0 @ c).real, np.vdot(c, z3 @ c).real, np.vdot(c, H @ c).real]
        assert abs(rows[k, 0] - t) < 1e-12
        assert np.allclose(means[k], want, rtol=0, atol=1e-9), (k, means[k], want)
    # the magnetisations move: the couplings were applied
    assert np.ptp(means[:, 0]) > 0.1


# The reference's demo configuration (configs/cpmg8z.toml), written out here
# key for key: L = 10, random initial state, dipolar geometry, field_disorder
# 100, a measure at t = 0, then 20 x [0.05 free, 1.07 pi pulse about x of
# hx = 10700, 0.05 free, measure].
CPMG8Z = """
[system]
spins = 10
initial_state = "random"
[field]
hx = 0.0
hy = 0.0
hz = 0.0
[ensemble]
realizations = 180
seed = 20240726
dipolar_geometry = true
dipolar_scale = 1.0
field_disorder = 100.0
[measurement]
reference_axis = "x"
observables = ["corr:x", "norm"]
[numerics]
epsilon = 1e-7
[[step]]
kind = "measure"
[[step]]
kind = "repeat"
count = 20
body = [
  { kind = "evolve", tau = 0.05, steps = 1 },
  { kind = "evolve", tau = 0.000314159265, steps = 1, hx = 10700.0 },
  { kind = "evolve", tau = 0.05, steps = 1 },
  { kind = "measure" },
]
"""


@pytest.mark.gpu
def test_demo_run_reproduces_the_documented_csv_row(tmp_path):
    """`spinsim run cpmg8z.toml --realizations 20` on the device: the t = 0
    row is, character for character, the one the reference documents
    (docs/configuration.md:229-231), and every later row agrees with the
    oracle's ensemble to the CSV's digits (1e-10 between device and oracle)."""
    from test_oracle import CPMG8Z_T0_ROW
    cfg = write(tmp_path, "cpmg8z.toml", CPMG8Z)
    r = spinsim("run", cfg, "--realizations", "20", "--threads", "4")
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[0].startswith("time,corr_x_re_mean,corr_x_re_sd,corr_x_re_stderr,")
    assert lines[1].startswith(CPMG8Z_T0_ROW + ",")
    header, rows = read_csv(r.stdout)
    p = O.Program().measure().repeat(20).evolve(0.05, 1)
    p.evolve(0.000314159265, 1, fields=np.tile([10700.0, 0.0, 0.0], (10, 1))).evolve(0.05, 1).measure().end()
    cfg_o = O.Config(L=10, program=p, seed=20240726, field_disorder=100.0,
                     observables=((O.OBS_CORR_X, 0), (O.OBS_NORM, 0)), reference_axis=0, epsilon=1e-7)
    times, _, mean, sd, se = O.run_ensemble(cfg_o, 20, threads=8)
    assert rows.shape == (21, 1 + 3 * mean.shape[1])
    assert np.allclose(rows[:, 0], times, rtol=1e-9, atol=0)   # the time column has 10 digits
    for k, want in enumerate((mean, sd, se)):
        assert np.allclose(rows[:, 1 + k::3], want, rtol=1e-9, atol=1e-10), k
