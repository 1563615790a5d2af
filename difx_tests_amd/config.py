"""Correlator config files: the reference's grammar (main.c:14-129).

Line oriented, split on ' ' and '\\n' only.  Keywords (case-insensitive):
NBIT COMPLEX NPOL NCHAN LO BANDWIDTH NUMFFTS NANT; NANT n is followed by n
lines `NAME FILE c0 c1 c2 OFFSET` (delay polynomial in seconds with t in FFT
units, file start offset in seconds; loadConfig.h:17-31).  Defaults NPOL=2,
COMPLEX=0, NBIT=2.  An empty line is an error (main.c:37-40); an unknown
keyword only warns (main.c:123-125).

Extensions read by the uneven-subband host program (fxbatch), which the
reference would warn about and skip:
  SUBBAND nchan bandwidth lo numffts antennas   (antennas: comma-separated
                                                 indices into the NANT list,
                                                 or ALL)
  NSUBINT n                                     (subints to process)
  PREC F32|F16|F8|F32V                          (X-engine precision)
and an antenna FILE may hold one %d, replaced by the subband index.
"""
from __future__ import annotations

import sys
from dataclasses import dataclass, field


class ConfigError(ValueError):
    pass


@dataclass
class AntennaEntry:
    name: str
    file: str
    delay: tuple[float, float, float]
    fileoffset: float


@dataclass
class SubbandEntry:
    nchan: int
    bandwidth: float
    lo: float
    numffts: int
    antennas: list[int] | None = None        # None: every antenna


@dataclass
class Config:
    nbit: int = 2
    npol: int = 2
    iscomplex: int = 0
    nchan: int = 0
    lo: float = 0.0
    bandwidth: float = 0.0
    numffts: int = 0
    antennas: list[AntennaEntry] = field(default_factory=list)
    subbands: list[SubbandEntry] = field(default_factory=list)
    nsubint: int = 1
    prec: str = "F32"

    @property
    def nant(self) -> int:
        return len(self.antennas)

    @property
    def subintbytes(self) -> int:
        """allocDataHost.c:21."""
        cfactor = 1 if self.iscomplex else 2
        return self.nchan * cfactor * (self.numffts + 1) * self.nbit // 8 * self.npol


_INT_KEYS = {"COMPLEX": "iscomplex", "NBIT": "nbit", "NPOL": "npol", "NCHAN": "nchan", "NUMFFTS": "numffts"}
_FLOAT_KEYS = {"LO": "lo", "BANDWIDTH": "bandwidth"}


def _tokens(line: str) -> list[str]:
    return [t for t in line.replace("\n", " ").split(" ") if t]


def _c_int(tok: str) -> int:
    """sscanf("%d") semantics: leading integer prefix."""
    s = tok.strip()
    i = 1 if s[:1] in "+-" else 0
    while i < len(s) and s[i].isdigit():
        i += 1
    if i == 0 or s[:i] in "+-":
        raise ConfigError(f"not an integer: {tok!r}")
    return int(s[:i])


def parse_config_text(text: str, warn=sys.stderr) -> Config:
    cfg = Config()
    ant_to_read = 0
    for line in text.splitlines(keepends=True):
        toks = _tokens(line)
        if not toks:
            raise ConfigError(f'Could not parse "{line}"')
        if ant_to_read:
            if len(toks) < 2:
                raise ConfigError(f"Insufficient data for antenna {len(cfg.antennas)}")
            if len(toks) < 6:
                raise ConfigError(f"Insufficient data for antenna {len(cfg.antennas)}")
            d = tuple(float(t) for t in toks[2:5])
            cfg.antennas.append(AntennaEntry(toks[0], toks[1], d, float(toks[5])))
            ant_to_read -= 1
            continue
        key = toks[0].upper()
        if key == "SUBBAND":
            if len(toks) < 6:
                raise ConfigError("Insufficient data for SUBBAND")
            ants = None if toks[5].upper() == "ALL" else [_c_int(t) for t in toks[5].split(",") if t]
            cfg.subbands.append(SubbandEntry(_c_int(toks[1]), float(toks[2]), float(toks[3]), _c_int(toks[4]), ants))
            continue
        if key == "NSUBINT":
            if len(toks) < 2:
                raise ConfigError("Insufficient data for NSUBINT")
            cfg.nsubint = _c_int(toks[1])
            continue
        if key == "PREC":
            if len(toks) < 2 or toks[1].upper() not in ("F32", "F16", "F8", "F32V"):
                raise ConfigError("PREC must be F32, F16, F8 or F32V")
            cfg.prec = toks[1].upper()
            continue
        if key in _INT_KEYS or key in _FLOAT_KEYS or key == "NANT":
            if len(toks) < 2:
                raise ConfigError(f"Insufficient data for {key}")
            if key == "NANT":
                n = _c_int(toks[1])
                if n < 1:
                    raise ConfigError("NANT must be positive")
                ant_to_read = n
                cfg.antennas = []
            elif key in _INT_KEYS:
                setattr(cfg, _INT_KEYS[key], _c_int(toks[1]))
            else:
                setattr(cfg, _FLOAT_KEYS[key], float(toks[1]))
        else:
            if warn is not None:
                print(f'Error: Unknown keyword "{toks[0]}"', file=warn)
    if ant_to_read:
        raise ConfigError(f"Insufficient data for antenna {len(cfg.antennas)}")
    return cfg


def parse_config(path: str, warn=sys.stderr) -> Config:
    with open(path) as f:
        return parse_config_text(f.read(), warn)


def format_config(cfg: Config) -> str:
    lines = [f"NBIT {cfg.nbit}", f"COMPLEX {cfg.iscomplex}", f"NCHAN {cfg.nchan}", f"LO {cfg.lo:.17g}",
             f"BANDWIDTH {cfg.bandwidth:.17g}", f"NUMFFTS {cfg.numffts}", f"NPOL {cfg.npol}",
             f"NANT {cfg.nant}"]
    for a in cfg.antennas:
        lines.append(f"{a.name} {a.file} {a.delay[0]:.17g} {a.delay[1]:.17g} {a.delay[2]:.17g} {a.fileoffset:.17g}")
    for sb in cfg.subbands:
        ants = "ALL" if sb.antennas is None else ",".join(str(a) for a in sb.antennas)
        lines.append(f"SUBBAND {sb.nchan} {sb.bandwidth:.17g} {sb.lo:.17g} {sb.numffts} {ants}")
    if cfg.subbands or cfg.nsubint != 1:
        lines.append(f"NSUBINT {cfg.nsubint}")
    if cfg.prec != "F32":
        lines.append(f"PREC {cfg.prec}")
    return "\n".join(lines) + "\n"
