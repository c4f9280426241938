"""fennol_md-style input decks (``.fnl``) run on the device MD driver.

``python -m fennol_amd.deck input.fnl`` plays the role of FeNNol's ``fennol_md``
(md/dynamic.py:52-150) for the ANI path: it parses the deck with the
reference's input syntax (utils/input_parser.py:144-253), reads the last frame
of the xyz/arc file (utils/io.py, md/initial.py:45-110), loads the ``.fnx``
model, and runs :class:`~fennol_amd.md.LangevinMD` with the deck's time step,
thermostat, temperature, friction, neighbour-list skin and periodic cell,
printing the energies every ``nprint`` steps and writing the trajectory every
``tdump``.

Keywords read (names and defaults as fennol_md's): model_file, xyz_input{file,
indexed, has_comment_line}, cell, minimum_image, nsteps, dt, temperature,
thermostat (LGV | NVE), gamma, random_seed, nblist_skin, nblist_stride,
nblist_warmup_time, nprint, tdump, traj_format (xyz | arc), system_name,
energy_unit (printing).  Keywords of other fennol_md features (barostats,
PIMD, QTB, colvars, ...) are rejected with an error rather than ignored.
"""

from __future__ import annotations

import math
import re
import sys
import time
from pathlib import Path

import numpy as np

from . import units

_SEPARATORS = " |,|=|\t|\n"
_COMMENTS = ("#", "!")
_TRUE = ("true", "yes", ".true.")
_FALSE = ("false", "no", ".false.")

# value of one MD unit (Angstrom, ps, kcal/mol, K) in the named unit
_TIME = {"FS": 1000.0, "PS": 1.0, "NS": 1.0e-3, "AU_T": 1.0 / units.PS}
_LENGTH = {"A": 1.0, "ANG": 1.0, "ANGSTROM": 1.0, "BOHR": 1.0 / units.ANGSTROM, "AU": 1.0 / units.ANGSTROM,
           "NM": 0.1}
_ENERGY = {"KCALPERMOL": 1.0, "KCAL/MOL": 1.0, "EV": units.EV / units.KCALPERMOL,
           "HA": 1.0 / units.KCALPERMOL, "HARTREE": 1.0 / units.KCALPERMOL, "KJPERMOL": 4.184, "KJ/MOL": 4.184}
_OTHER = {"K": 1.0, "THZ": 1.0, "PS^-1": 1.0, "PS^{-1}": 1.0, "FS^-1": 1.0e-3, "FS^{-1}": 1.0e-3}

# fennol_md keywords of features this build does not run
_UNSUPPORTED = ("barostat", "nbeads", "nreplicas", "colvars", "qtb", "thermostat_file", "restart_file",
                "additional_keys", "deuterate", "hmr", "ir_spectrum")


def unit_multiplier(unit: str) -> float:
    """Value of one MD unit in ``unit`` (the reference's us.get_multiplier for the simple units
    a deck uses: time, length, energy, temperature, frequency)."""
    u = unit.strip().upper()
    for table in (_TIME, _LENGTH, _ENERGY, _OTHER):
        if u in table:
            return table[u]
    raise ValueError(f"unit [{unit}] is not supported in decks here")


class Deck(dict):
    """Parsed deck: keys are lower-case paths joined with '/' (``xyz_input/file``)."""

    def get(self, key, default=None):
        return super().get(key.lower(), default)


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


def parse_deck(path) -> Deck:
    """Parse a ``.fnl`` deck (input_parser.py:144-233): '#'/'!' comments, separators space,
    comma, '=' and tab, ``name{ ... }`` sections, ``key[unit]`` converted to MD units, a key
    alone is True, several values make a list."""
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
        if joined.endswith("{"):
            section.append(joined[:-1].lower())
            continue
        if joined[0] in "}/" or "&end" in joined:
            section.pop()
            continue
        key = words[0].lower()
        mult = None
        k = max(key.find("["), key.find("{"))
        if k == 0:
            raise ValueError(f"field {words[0]!r} must not start with a unit")
        if k > 0:
            close = "]" if key[k] == "[" else "}"
            if not key.endswith(close):
                raise ValueError(f"wrong unit specification in field {words[0]!r}")
            mult = unit_multiplier(key[k + 1:-1]) if len(key) - k > 2 else 1.0
            key = key[:k]
        if len(words) == 1:
            val = True
        elif len(words) == 2:
            val = _true_type(words[1], mult)
        else:
            val = [_true_type(w, mult) for w in words[1:]]
        deck["/".join(section + [key])] = val
    return deck


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
    """Run the deck at ``path``; returns a summary (steps, energies, timings, files)."""
    from .fennix import FENNIX
    from .md import LangevinMD

    path = Path(path)
    base = workdir if workdir is not None else path.parent
    deck = parse_deck(path)
    bad = [k for k in deck if k.split("/")[0] in _UNSUPPORTED]
    if bad:
        raise NotImplementedError(f"deck keywords not supported by this build: {bad}")
    model_file = base / str(deck.get("model_file", "")).strip()
    if not model_file.is_file():
        raise FileNotFoundError(f"model file {model_file} not found")
    xyz_file = base / str(deck.get("xyz_input/file", str(deck.get("system_name", "system")) + ".xyz")).strip()
    if not xyz_file.is_file():
        raise FileNotFoundError(f"xyz file {xyz_file} not found")
    symbols, coords = read_last_frame(xyz_file, indexed=bool(deck.get("xyz_input/indexed", False)),
                                      has_comment_line=bool(deck.get("xyz_input/has_comment_line", True)))
    name = str(deck.get("system_name", xyz_file.stem)).strip()
    species = np.array([units.SYMBOL_TO_Z[s.capitalize()] for s in symbols], np.int32)
    cell = _cell(deck)
    model = FENNIX.load(model_file, precision=precision)
    system = {"species": species, "coordinates": coords}
    if cell is not None:
        system["cells"] = cell[None]
    nsteps = int(deck.get("nsteps"))
    dt_fs = float(deck.get("dt")) * 1000.0              # ps -> fs
    temperature = float(deck.get("temperature", 300.0))
    thermostat = str(deck.get("thermostat", "LGV")).upper()
    gamma_thz = float(deck.get("gamma", 1.0))
    seed = int(deck.get("random_seed", np.random.randint(0, 2 ** 31 - 1)))
    skin = float(deck.get("nblist_skin", -1.0))
    md = LangevinMD(model.potential, system, dt_fs=dt_fs, temperature=temperature, gamma_thz=gamma_thz,
                    seed=seed, thermostat=thermostat, energy_unit=model.energy_unit,
                    nblist_skin=skin, nblist_stride=int(deck.get("nblist_stride", -1)),
                    nblist_warmup_fs=float(deck.get("nblist_warmup_time", -1.0)) * 1000.0,
                    minimum_image=bool(deck.get("minimum_image", True)))
    nprint = int(deck.get("nprint", 10))
    tdump_ps = float(deck.get("tdump", 1.0))
    ndump = max(1, int(round(tdump_ps / (dt_fs / 1000.0))))
    fmt = str(deck.get("traj_format", "arc")).lower()
    if fmt not in ("xyz", "arc"):
        raise NotImplementedError(f"traj_format {fmt}")
    traj = base / f"{name}.traj.{fmt}"
    e_unit = str(deck.get("energy_unit", "kcal/mol"))
    e_mult = unit_multiplier(e_unit)
    n = len(species)
    print(f"# model_file: {model_file}\n# system: {name}, {n} atoms, thermostat {thermostat}, "
          f"dt {dt_fs} fs, T {temperature} K, skin {skin} A (rebuild every {md.nblist_stride} steps)", file=out)
    print(f"#{'Step':>9} {'Time[ps]':>10} {'Etot':>13} {'Epot':>13} {'Ekin':>13} {'Temp[K]':>9} "
          f"{'ns/day':>9}   (energies per atom, {e_unit})", file=out)
    rows = []
    t0 = time.perf_counter()
    tlast, slast = t0, 0
    with open(traj, "w") as fh:
        for istep in range(1, nsteps + 1):
            md.step()
            if istep % ndump == 0:
                _write_frame(fh, fmt, symbols, md.x.cpu().numpy(), cell, f"step {istep}")
            if istep % nprint == 0 or istep == nsteps:
                ek = float(np.trace(md.kinetic_tensor()[0]))
                ep = float(md.energies()[0]) / md.model_energy_unit
                temp = 2.0 * ek / (3.0 * n * units.K_B)
                now = time.perf_counter()
                nsday = (istep - slast) * dt_fs * 1e-6 / max(now - tlast, 1e-12) * 86400.0
                tlast, slast = now, istep
                rows.append((istep, ep + ek, ep, ek, temp))
                print(f"{istep:10d} {istep * dt_fs / 1000.0:10.4f} {(ep + ek) * e_mult / n:13.6f} "
                      f"{ep * e_mult / n:13.6f} {ek * e_mult / n:13.6f} {temp:9.2f} {nsday:9.2f}", file=out)
                if not math.isfinite(ep):
                    raise FloatingPointError(f"step {istep}: non-finite energy")
    md.verify()
    wall = time.perf_counter() - t0
    print(f"# Run done in {wall:.2f} s ({nsteps / wall:.1f} step/s)", file=out)
    return {"steps": nsteps, "rows": rows, "wall_s": wall, "trajectory": str(traj), "md": md}


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m fennol_amd.deck input.fnl", file=sys.stderr)
        return 2
    run_deck(argv[0])
    return 0


if __name__ == "__main__":
    sys.exit(main())
