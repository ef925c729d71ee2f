"""Legacy .dat pulse-sequence scripts to run configurations (TOML).

The reference migrates its archived Fortran-era inputs with
tools/dat2toml.py (docs/legacy.md:103-128); this module reads the same two
nested formats and writes a configuration that `spinsim validate` and
`spinsim run` (spinsim_amd/host/apps/spinsim.cpp) accept:

  * a script: one `@file.dat` line per step, with `@CYCLE n` ... `@ENDCYCLE`
    around the repeated block;
  * step files: keyword lines, most followed by a value line, blocks closed
    by `END`.

The step-file keywords and what they carry (tools/dat2toml.py:62-140):
  HEISENBERG INTERACTION  rows `i, j, Jx, Jy, Jz` until END; the largest
                          index sets the system size, Jz the dipolar scale
  SPIN : k                next line: Hx0, Hx1, Omega_x, Hy0, ... ; only the
                          static parts (columns 0, 3, 6) are used
  NUMBER OF FIELD STEPS   evolve steps; 0 marks the setup file
  TIME STEP               tau
  NUMBER OF INTERMEDIATE TIME STEPS   multiplies tau (it does not subdivide)
  INITIAL STATE: ...      SPIN UP / SPIN DOWN / RANDOM
  SPECIAL PROPAGATION, GROUND STATE   not representable: noted in the output
  IMAGINARY TIME PROPAGATION          the step evolves in imaginary time

Three settings lived in Fortran source, not in any data file (the measured
correlator, the realisation count and the disorder strength); the output
carries the reference converter's defaults for them, each marked TODO.

Where this differs from tools/dat2toml.py, by design:
  * every evolve step writes all three field components, so a component the
    step file sets to zero does not fall back to a non-zero baseline;
  * fields that differ between spins become per-spin arrays when every spin
    has a SPIN row (otherwise spin 1's values, with a warning, as there);
  * steps after @ENDCYCLE are kept, after the repeat;
  * IMAGINARY TIME PROPAGATION sets mode = "imaginary" on its step.

Usage: python -m spinsim_amd.legacy SCRIPT.dat [-o OUT.toml]
"""
from __future__ import annotations

import argparse
import pathlib
import re
import sys
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple


class LegacyError(ValueError):
    """A script or step file that cannot be converted."""


@dataclass
class LegacyStep:
    """What one step file defines."""

    name: str
    nspins: int = 0
    fields: Dict[int, Tuple[float, float, float]] = field(default_factory=dict)   # 1-based spin
    tau: float = 0.0
    intermediate: int = 1
    nsteps: int = 0
    jz_scale: Optional[float] = None
    initial_state: Optional[str] = None
    imaginary: bool = False
    notes: List[str] = field(default_factory=list)

    @property
    def duration(self) -> float:
        return self.tau * self.intermediate


@dataclass
class LegacyScript:
    prologue: List[LegacyStep]
    body: List[LegacyStep]
    count: int
    epilogue: List[LegacyStep]

    @property
    def steps(self) -> List[LegacyStep]:
        return self.prologue + self.body + self.epilogue


_SPLIT = re.compile(r"[,\s]+")


def _numbers(line: str) -> List[str]:
    return [p for p in _SPLIT.split(line.strip()) if p]


def read_step_file(path: pathlib.Path) -> LegacyStep:
    lines = path.read_text().splitlines()
    step = LegacyStep(name=path.name)

    def value_line(i: int) -> str:
        if i + 1 >= len(lines):
            raise LegacyError(f"{path.name}: '{lines[i].strip()}' at line {i + 1} has no value line")
        return lines[i + 1].strip()

    i = 0
    while i < len(lines):
        line = lines[i].strip()
        if line.startswith("HEISENBERG INTERACTION"):
            i += 1
            while i < len(lines) and not lines[i].strip().startswith("END"):
                row = _numbers(lines[i])
                if len(row) >= 5:
                    step.nspins = max(step.nspins, int(row[0]), int(row[1]))
                    step.jz_scale = float(row[4])
                i += 1
            continue
        if line.startswith("SPIN :"):
            spin = int(_numbers(line.split(":", 1)[1])[0])
            row = _numbers(value_line(i))
            if len(row) < 9:
                raise LegacyError(f"{path.name}: SPIN row at line {i + 2} has {len(row)} of 9 columns")
            step.fields[spin] = (float(row[0]), float(row[3]), float(row[6]))
            step.nspins = max(step.nspins, spin)
            i += 2
            continue
        if line.startswith("NUMBER OF FIELD STEPS"):
            step.nsteps = int(value_line(i))
            i += 2
            continue
        if line.startswith("NUMBER OF INTERMEDIATE TIME STEPS"):
            step.intermediate = int(value_line(i))
            i += 2
            continue
        if line.startswith("TIME STEP"):
            step.tau = float(value_line(i))
            i += 2
            continue
        if line.startswith("INITIAL STATE"):
            for keyword, kind in (("SPIN UP", "up"), ("SPIN DOWN", "down"), ("RANDOM", "random")):
                if keyword in line:
                    step.initial_state = kind
            if "GROUND STATE" in line:
                step.notes.append(f"{path.name}: INITIAL STATE: GROUND STATE is not supported; random is used")
        elif "SPECIAL PROPAGATION" in line:
            step.notes.append(f"{path.name}: SPECIAL PROPAGATION (an ideal gate) needs a "
                              'hand-written [[step]] of kind = "pulse"')
        elif "IMAGINARY TIME PROPAGATION" in line:
            step.imaginary = True
            step.notes.append(f'{path.name}: imaginary-time step, written with mode = "imaginary"')
        i += 1
    return step


def read_script(path: pathlib.Path) -> LegacyScript:
    path = pathlib.Path(path)
    parts: List[List[LegacyStep]] = [[]]
    count = 1
    in_cycle = False
    for raw in path.read_text().splitlines():
        line = raw.strip()
        if not line:
            continue
        if line.startswith("@CYCLE"):
            if in_cycle or len(parts) > 1:
                raise LegacyError(f"{path.name}: a second @CYCLE is not representable")
            tail = line[len("@CYCLE"):].strip()
            count = int(tail) if tail else 1
            in_cycle = True
            parts.append([])
        elif line.startswith("@ENDCYCLE"):
            if not in_cycle:
                raise LegacyError(f"{path.name}: @ENDCYCLE without @CYCLE")
            in_cycle = False
            parts.append([])
        elif line.startswith("@"):
            target = (path.parent / line[1:].strip()).resolve()
            if not target.exists():
                raise LegacyError(f"{path.name}: {line} does not exist")
            parts[-1].append(read_step_file(target))
    if not any(parts):
        raise LegacyError(f"{path.name}: the script includes no step files")
    prologue = parts[0]
    body = parts[1] if len(parts) > 1 else []
    epilogue = parts[2] if len(parts) > 2 else []
    return LegacyScript(prologue, body, count if len(parts) > 1 else 1, epilogue)


def _number(x: float) -> str:
    return repr(float(x))


def _field_values(step: LegacyStep, nspins: int, notes: List[str]) -> Tuple[str, str, str]:
    """hx, hy, hz as TOML values: a scalar for a component every spin shares,
    else a per-spin array."""
    if not step.fields:
        return "0.0", "0.0", "0.0"
    complete = all(s in step.fields for s in range(1, nspins + 1))
    out = []
    for a in range(3):
        values = [v[a] for v in step.fields.values()]
        if all(x == values[0] for x in values):
            out.append(_number(values[0]))
        elif complete:
            out.append("[" + ", ".join(_number(step.fields[s][a]) for s in range(1, nspins + 1)) + "]")
        else:
            first = step.fields[min(step.fields)]
            notes.append(f"{step.name}: fields differ between spins but not every spin has a SPIN row; "
                         "spin 1's were used for all, write the arrays by hand")
            return tuple(_number(c) for c in first)
    return tuple(out)


def _evolve(step: LegacyStep, nspins: int, notes: List[str]) -> List[str]:
    """The `key = value` entries of an evolve step."""
    hx, hy, hz = _field_values(step, nspins, notes)
    parts = ['kind = "evolve"', f"tau = {_number(step.duration)}", f"steps = {step.nsteps}",
             f"hx = {hx}", f"hy = {hy}", f"hz = {hz}"]
    if step.imaginary:
        parts.append('mode = "imaginary"')
    return parts


def to_toml(script: LegacyScript, source: str) -> str:
    steps = script.steps
    nspins = max(s.nspins for s in steps)
    if nspins < 1:
        raise LegacyError("no step file names a spin")
    jz = next((s.jz_scale for s in steps if s.jz_scale), 1.0)
    initial = next((s.initial_state for s in steps if s.initial_state), "random")
    notes: List[str] = [n for s in steps for n in s.notes]
    setup = next((s for s in steps if s.nsteps == 0), None)
    base = _field_values(setup, nspins, notes) if setup else ("0.0", "0.0", "0.0")

    def block(seq: List[LegacyStep]) -> List[str]:
        out = []
        for s in seq:
            if s.nsteps == 0:
                out.append(f"# {s.name}: no field steps (setup only)")
                continue
            out += ["[[step]]"] + _evolve(s, nspins, notes) + [""]
        return out

    seq: List[str] = ["[[step]]", 'kind = "measure"', ""]
    seq += block(script.prologue)
    if script.body:
        seq += ["[[step]]", 'kind = "repeat"', f"count = {script.count}", "body = ["]
        for s in script.body:
            seq.append(f"  # {s.name}: no field steps (setup only)" if s.nsteps == 0
                       else "  { " + ", ".join(_evolve(s, nspins, notes)) + f" }},  # {s.name}")
        seq += ['  { kind = "measure" },', "]", ""]
    seq += block(script.epilogue)

    head = [f"# Converted from {source} by spinsim_amd.legacy.", "#",
            "# The TODOs mark settings the legacy code kept in Fortran source, not",
            "# in any data file; no converter can recover them."]
    head += ["#"] + [f"# WARNING: {n}" for n in notes] if notes else []
    head += ["", "[system]", f"spins = {nspins}", f'initial_state = "{initial}"', "",
             "[field]", f"hx = {base[0]}", f"hy = {base[1]}", f"hz = {base[2]}", "",
             "[ensemble]", "# TODO: the legacy realisation count (Nit) was a Fortran parameter",
             "realizations = 180", "seed = 1", "dipolar_geometry = true", f"dipolar_scale = {_number(jz)}",
             "# TODO: the legacy disorder strength (HzScale) was a Fortran parameter",
             "field_disorder = 100.0", "",
             "[measurement]", "# TODO: the measured correlator was chosen in Fortran source",
             '# (sum sigma_x for the CPMG runs, sum sigma_z for the Rabi runs)',
             'reference_axis = "z"', 'observables = ["corr:z", "norm"]', "",
             "[numerics]", "epsilon = 1e-7", ""]
    return "\n".join(head + seq)


def convert(path) -> str:
    """TOML text for the legacy @-script at `path`."""
    return to_toml(read_script(pathlib.Path(path)), str(path))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m spinsim_amd.legacy",
                                 description="Convert a legacy .dat @-script into a run configuration.")
    ap.add_argument("script", type=pathlib.Path)
    ap.add_argument("-o", "--out", type=pathlib.Path, help="write here instead of standard output")
    args = ap.parse_args(argv)
    try:
        text = convert(args.script)
    except (LegacyError, OSError, ValueError) as e:
        print(f"legacy: {e}", file=sys.stderr)
        return 1
    if args.out:
        args.out.write_text(text)
    else:
        sys.stdout.write(text)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
