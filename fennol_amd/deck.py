"""fennol_md-style input decks (``.fnl``) run on the device MD driver.

``python -m fennol_amd.deck input.fnl`` plays the role of FeNNol's ``fennol_md``
(md/dynamic.py:52-652) for the ANI path: it parses the deck with the
reference's input syntax (utils/input_parser.py:144-279, units converted
through the MD unit system as ``us.get_multiplier`` does), reads the last
frame of the xyz/arc file (utils/io.py, md/initial.py:45-110), loads the
``.fnx`` model, and runs :class:`~fennol_amd.md.LangevinMD` with the deck's
time step, thermostat, temperature, friction, neighbour-list skin and
periodic cell, printing the energies (and, with ``estimate_pressure``, the
pressure) every ``nprint`` steps and writing the trajectory every ``tdump``.

Keywords acted on (names and defaults as fennol_md's): model_file,
xyz_input{file, indexed, has_comment_line}, system_name, cell,
minimum_image, wrap_box, estimate_pressure, pressure_unit, nsteps, dt,
temperature, thermostat (LGV | LANGEVIN | NVE), gamma, random_seed,
nblist_skin, nblist_stride, nblist_warmup_time, nprint, tdump, traj_format
(xyz | arc), energy_unit, per_atom_energy.  Keywords that change nothing on
this path are accepted (device, matmul_prec, print_timings,
double_precision -- positions and velocities are FP64 here anyway --,
nsummary, nblist_verbose).  Features this build does not run raise when they
are *active* (a thermostat other than LGV/NVE, a barostat, path integrals,
replicas, colvars, IR spectra, deuteration, HMR, minimisation, restarts); a
section that only configures an inactive feature (``qtb{}`` without a QTB
thermostat) is parsed and left alone.
"""

from __future__ import annotations

import math
import re
import sys
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import units

_SEPARATORS = " |,|=|\t|\n"
_COMMENTS = ("#", "!")
_TRUE = ("true", "yes", ".true.")
_FALSE = ("false", "no", ".false.")

_THERMOSTATS = ("LGV", "LANGEVIN", "NVE", "NONE")


def unit_multiplier(unit: str) -> float:
    """Value of one MD unit (Angstrom, ps, kcal/mol) in ``unit`` (us.get_multiplier)."""
    return units.us.get_multiplier(unit)


class Deck(dict):
    """Parsed deck: keys are lower-case paths joined with '/' (``xyz_input/file``)."""

    def get(self, key, default=None):
        return super().get(key.lower(), default)

    def section(self, name: str) -> dict:
        p = name.lower() + "/"
        return {k[len(p):]: v for k, v in self.items() if k.startswith(p)}


def _true_type(word: str, mult: float | None):
    if mult is not None:
        return float(word) / mult
    for conv in (int, float):
        try:
            return conv(word)
        except ValueError:
            pass
    if word.lower() in _TRUE:
        return True
    if word.lower() in _FALSE:
        return False
    return word


def _key_unit(word: str):
    """``key[unit]`` / ``key{unit}`` -> (key, multiplier or None) (input_parser.py:255-279)."""
    k = max(word.find("["), word.find("{"))
    if k < 0:
        return word, None
    if k == 0:
        raise ValueError(f"field {word!r} must not start with '{{' or '['")
    close = "]" if word[k] == "[" else "}"
    if not word.endswith(close):
        raise ValueError(f"wrong unit specification in field {word!r}")
    if len(word) - k <= 2:
        return word[:k], 1.0
    return word[:k], unit_multiplier(word[k + 1:-1])


def parse_deck(path) -> Deck:
    """Parse a ``.fnl`` deck (input_parser.py:144-233): '#'/'!' comments, separators space,
    comma, '=' and tab; sections ``name{ ... }`` or ``&name ... &end`` (closed by a line
    starting with '}' or '/', or holding '&end'), ``name{}`` an empty section; ``key[unit]``
    converted to MD units; a key alone is True; several values make a list."""
    deck = Deck()
    section: list[str] = []
    for raw in Path(path).read_text().splitlines():
        line = raw
        for c in _COMMENTS:
            k = line.find(c)
            if k >= 0:
                line = line[:k]
        words = [w for w in re.split(_SEPARATORS, line.strip()) if w]
        if not words:
            continue
        joined = "".join(words)
        if joined.endswith("{}"):
            deck["/".join(section + [joined[:-2].lower()])] = Deck()
            continue
        if joined.endswith("{"):
            section.append(joined[:-1].lower())
            continue
        if joined.startswith("&") and "&end" not in joined.lower():
            section.append(joined[1:].lower())
            continue
        if joined[0] in "}/" or "&end" in joined.lower():
            if not section:
                raise ValueError(f"{path}: section closed that was never opened: {raw.strip()!r}")
            section.pop()
            continue
        key, mult = _key_unit(words[0].lower())
        if len(words) == 1:
            val = True
        elif len(words) == 2:
            val = _true_type(words[1], mult)
        else:
            val = [_true_type(w, mult) for w in words[1:]]
        deck["/".join(section + [key])] = val
    if section:
        raise ValueError(f"{path}: section {'/'.join(section)!r} is not closed")
    return deck


def _flag(deck: Deck, key: str, default=False) -> bool:
    v = deck.get(key, default)
    if isinstance(v, str):
        return v.lower() in _TRUE
    return bool(v)


@dataclass
class DeckPlan:
    """What a deck asks for, resolved to MD units and checked against this build."""

    model_file: Path
    xyz_file: Path
    indexed: bool
    has_comment_line: bool
    system_name: str | None
    cell: np.ndarray | None
    minimum_image: bool
    wrap_box: bool
    estimate_pressure: bool
    pressure_unit: str
    nsteps: int
    dt_fs: float
    temperature: float
    thermostat: str
    gamma_thz: float
    seed: int | None
    nblist_skin: float
    nblist_stride: int
    nblist_warmup_fs: float
    nprint: int
    tdump_ps: float
    traj_format: str
    energy_unit: str
    per_atom_energy: bool


def _cell(deck: Deck):
    c = deck.get("cell")
    if c is None:
        return None
    c = np.asarray(c, np.float64).reshape(-1)
    if c.size == 3:
        return np.diag(c)
    if c.size == 9:
        return c.reshape(3, 3)  # rows = lattice vectors
    raise ValueError("cell: 3 lengths or 9 components (ax ay az bx by bz cx cy cz)")


def plan_deck(deck: Deck, base: Path) -> DeckPlan:
    """Resolve and check a parsed deck (files are not opened).  Raises NotImplementedError for
    active features outside this build and ValueError for malformed values."""
    thermostat = str(deck.get("thermostat", "LGV")).strip().upper()
    if thermostat not in _THERMOSTATS:
        raise NotImplementedError(f"thermostat {thermostat}: only LGV (Langevin) and NVE run here")
    active = []
    barostat = deck.get("barostat")
    if barostat is not None and str(barostat).strip().upper() not in ("NONE", "NVT", "FALSE"):
        active.append(f"barostat {barostat}")
    for key in ("nbeads", "nreplicas", "colvars", "ir_spectrum", "deuterate", "restart_file",
                "xyz_input/crystal", "xyz_input/minimize", "topology", "etot_ensemble_key"):
        v = deck.get(key)
        if v not in (None, False, 0) and not (isinstance(v, dict) and not v):
            active.append(key)
    if deck.section("colvars"):
        active.append("colvars{}")
    if int(deck.get("hmr", 0) or 0) > 0:
        active.append("hmr")
    if deck.get("additional_keys"):
        active.append("additional_keys")
    if deck.get("total_charge") not in (None, 0, 0.0):
        active.append("total_charge (the ANI path has no charge input)")
    if active:
        raise NotImplementedError(f"deck features not supported by this build: {active}")
    if deck.get("nsteps") is None or deck.get("dt") is None:
        raise ValueError("the deck must give nsteps and dt")
    model = deck.get("model_file")
    if model is None:
        raise ValueError("the deck must give model_file")
    name = deck.get("system_name")
    xyz = deck.get("xyz_input/file", (str(name).strip() if name else "system") + ".xyz")
    traj_format = str(deck.get("traj_format", "arc")).lower()
    if traj_format not in ("xyz", "arc"):
        raise NotImplementedError(f"traj_format {traj_format}: xyz and arc are written here")
    cell = _cell(deck)
    seed = deck.get("random_seed")
    return DeckPlan(
        model_file=base / str(model).strip(), xyz_file=base / str(xyz).strip(),
        indexed=_flag(deck, "xyz_input/indexed", False),
        has_comment_line=_flag(deck, "xyz_input/has_comment_line", True),
        system_name=None if name is None else str(name).strip(), cell=cell,
        minimum_image=_flag(deck, "minimum_image", True), wrap_box=_flag(deck, "wrap_box", False),
        estimate_pressure=_flag(deck, "estimate_pressure", False),
        pressure_unit=str(deck.get("pressure_unit", "atm")).strip(),
        nsteps=int(deck.get("nsteps")), dt_fs=float(deck.get("dt")) * units.FS_PER_PS,
        temperature=float(deck.get("temperature", 300.0)), thermostat=thermostat,
        gamma_thz=float(deck.get("gamma", 1.0 / units.THZ)) * units.THZ,
        seed=None if seed is None else int(seed),
        nblist_skin=float(deck.get("nblist_skin", -1.0)), nblist_stride=int(deck.get("nblist_stride", -1)),
        nblist_warmup_fs=float(deck.get("nblist_warmup_time", -1.0)) * units.FS_PER_PS,
        nprint=int(deck.get("nprint", 10)), tdump_ps=float(deck.get("tdump", 1.0)),
        traj_format=traj_format, energy_unit=str(deck.get("energy_unit", "kcal/mol")).strip(),
        per_atom_energy=_flag(deck, "per_atom_energy", True))


def read_last_frame(path, indexed: bool = False, has_comment_line: bool = True):
    """Symbols and coordinates of the last frame of an xyz (or Tinker arc, ``indexed``) file."""
    lines = Path(path).read_text().splitlines()
    frames = []
    k = 0
    while k < len(lines):
        s = lines[k].strip()
        if not s or s.startswith("#"):
            k += 1
            continue
        nat = int(s.split()[0])
        k += 1 + (1 if has_comment_line else 0)
        if indexed and k < len(lines):  # Tinker arc files may carry a cell line before the atoms
            parts = lines[k].split()
            if len(parts) == 6 and not parts[0].isdigit():
                k += 1
        body = lines[k:k + nat]
        if len(body) < nat:
            raise ValueError(f"{path}: premature end of file")
        sym, xyz = [], []
        for ln in body:
            p = ln.split()
            if indexed:
                p = p[1:]
            sym.append(p[0])
            xyz.append([float(v) for v in p[1:4]])
        frames.append((sym, np.array(xyz)))
        k += nat
    if not frames:
        raise ValueError(f"{path}: no frame")
    return frames[-1]


def wrap_into_box(x: np.ndarray, cell: np.ndarray) -> np.ndarray:
    """Positions wrapped into the primary cell: q = x R, q -= floor(q), x = q C
    (md/utils.py:58-70 without atom groups)."""
    q = x @ np.linalg.inv(cell)
    return (q - np.floor(q)) @ cell


def _write_frame(fh, fmt: str, symbols, x, cell, comment: str) -> None:
    n = len(symbols)
    if fmt == "arc":
        fh.write(f"{n:6d}  {comment}\n")
        if cell is not None:
            a, b, c = (np.linalg.norm(v) for v in cell)
            al = math.degrees(math.acos(np.dot(cell[1], cell[2]) / (b * c)))
            be = math.degrees(math.acos(np.dot(cell[0], cell[2]) / (a * c)))
            ga = math.degrees(math.acos(np.dot(cell[0], cell[1]) / (a * b)))
            fh.write(f" {a:12.6f}{b:12.6f}{c:12.6f}{al:12.6f}{be:12.6f}{ga:12.6f}\n")
        for i, (s, r) in enumerate(zip(symbols, x)):
            fh.write(f"{i + 1:6d}  {s:3s}{r[0]:12.6f}{r[1]:12.6f}{r[2]:12.6f}\n")
    else:
        fh.write(f"{n}\n{comment}\n")
        for s, r in zip(symbols, x):
            fh.write(f"{s:3s}{r[0]:15.8f}{r[1]:15.8f}{r[2]:15.8f}\n")


def run_deck(path, out=sys.stdout, precision: str = "fp32", workdir: Path | None = None) -> dict:
    """Run the deck at ``path``; returns a summary (steps, printed rows, timings, files)."""
    from .fennix import FENNIX
    from .md import LangevinMD

    path = Path(path)
    base = workdir if workdir is not None else path.parent
    plan = plan_deck(parse_deck(path), base)
    if not plan.model_file.is_file():
        raise FileNotFoundError(f"model file {plan.model_file} not found")
    if not plan.xyz_file.is_file():
        raise FileNotFoundError(f"xyz file {plan.xyz_file} not found")
    symbols, coords = read_last_frame(plan.xyz_file, indexed=plan.indexed, has_comment_line=plan.has_comment_line)
    name = plan.system_name or plan.xyz_file.stem
    species = np.array([units.SYMBOL_TO_Z[s.capitalize()] for s in symbols], np.int32)
    cell = plan.cell
    if plan.estimate_pressure and cell is None:
        raise ValueError("estimate_pressure needs a periodic cell")
    model = FENNIX.load(plan.model_file, precision=precision)
    system = {"species": species, "coordinates": coords}
    if cell is not None:
        system["cells"] = cell[None]
    seed = plan.seed if plan.seed is not None else int(np.random.randint(0, 2 ** 31 - 1))
    md = LangevinMD(model.potential, system, dt_fs=plan.dt_fs, temperature=plan.temperature,
                    gamma_thz=plan.gamma_thz, seed=seed, thermostat=plan.thermostat,
                    energy_unit=model.energy_unit, nblist_skin=plan.nblist_skin, nblist_stride=plan.nblist_stride,
                    nblist_warmup_fs=plan.nblist_warmup_fs, minimum_image=plan.minimum_image)
    ndump = max(1, int(round(plan.tdump_ps / (plan.dt_fs / units.FS_PER_PS))))
    fmt = plan.traj_format
    traj = base / f"{name}.traj.{fmt}"
    e_unit = plan.energy_unit
    e_mult = unit_multiplier(e_unit)
    n = len(species)
    per = n if plan.per_atom_energy else 1
    p_mult = unit_multiplier(plan.pressure_unit) if plan.estimate_pressure else 1.0
    print(f"# model_file: {plan.model_file}\n# system: {name}, {n} atoms, thermostat {plan.thermostat}, "
          f"dt {plan.dt_fs} fs, T {plan.temperature} K, skin {plan.nblist_skin} A (rebuild every "
          f"{md.nblist_stride} steps)", file=out)
    head = (f"#{'Step':>9} {'Time[ps]':>10} {'Etot':>13} {'Epot':>13} {'Ekin':>13} {'Temp[K]':>9} "
            f"{'ns/day':>9}")
    if plan.estimate_pressure:
        head += f" {'Press[' + plan.pressure_unit + ']':>14}"
    print(head + f"   (energies{' per atom' if plan.per_atom_energy else ''}, {e_unit})", file=out)
    rows = []
    t0 = time.perf_counter()
    tlast, slast = t0, 0
    with open(traj, "w") as fh:
        for istep in range(1, plan.nsteps + 1):
            md.step()
            if istep % ndump == 0:
                x = md.x.cpu().numpy()
                if plan.wrap_box and cell is not None:
                    x = wrap_into_box(x, cell)
                _write_frame(fh, fmt, symbols, x, cell, f"step {istep}")
            if istep % plan.nprint == 0 or istep == plan.nsteps:
                ek = float(np.trace(md.kinetic_tensor()[0]))
                ep = float(md.energies()[0]) / md.model_energy_unit
                temp = 2.0 * ek / (3.0 * n * units.K_B)
                now = time.perf_counter()
                nsday = (istep - slast) * plan.dt_fs * 1e-6 / max(now - tlast, 1e-12) * 86400.0
                tlast, slast = now, istep
                row = [istep, ep + ek, ep, ek, temp]
                line = (f"{istep:10d} {istep * plan.dt_fs / 1000.0:10.4f} {(ep + ek) * e_mult / per:13.6f} "
                        f"{ep * e_mult / per:13.6f} {ek * e_mult / per:13.6f} {temp:9.2f} {nsday:9.2f}")
                if plan.estimate_pressure:
                    pres = float(md.pressure()[0]) * p_mult
                    row.append(pres)
                    line += f" {pres:14.4f}"
                rows.append(tuple(row))
                print(line, file=out)
                if not math.isfinite(ep):
                    raise FloatingPointError(f"step {istep}: non-finite energy")
    md.verify()
    wall = time.perf_counter() - t0
    print(f"# Run done in {wall:.2f} s ({plan.nsteps / wall:.1f} step/s)", file=out)
    return {"steps": plan.nsteps, "rows": rows, "wall_s": wall, "trajectory": str(traj), "md": md, "plan": plan}


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m fennol_amd.deck input.fnl", file=sys.stderr)
        return 2
    run_deck(argv[0])
    return 0


if __name__ == "__main__":
    sys.exit(main())
