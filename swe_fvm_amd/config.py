""".ini configuration and non-dimensional scaling: the Python twin of
include/swe_config.hpp (Parser, DimensionManager, Scales; reference
interfaces docs/ConfigParser_8h_source.html and
docs/DimensionManager_8h_source.html), and the construction of the analytic
tests from a configuration the way the reference's test classes read it
(examples/Tests.h:13-16,51,66-71,140-142).

File format: "key = value" lines under "[section]" headers; blank lines and
lines starting with ';' or '#' are skipped; everything is trimmed.
Scales: height h0, length l0, velocity c0 = sqrt(g h0), source c0 / l0, time
l0 / c0, with h0, l0 and g (default 1.0) read from [Dimensions].
"""
import enum
import math

from . import analytic


class ParserError(ValueError):
    pass


class DimensionError(ValueError):
    pass


class Parser:
    """Parser(filename) / Parser.from_string(text); Get(section, property)."""

    def __init__(self, filename=None, text=None):
        if text is None:
            try:
                with open(filename) as f:
                    text = f.read()
            except OSError as e:
                raise ParserError(f"cannot open {filename}") from e
        self.m_ini = {}
        section = ""
        for no, raw in enumerate(text.splitlines(), 1):
            line = raw.strip()
            if not line or line[0] in ";#":
                continue
            if line[0] == "[":
                if line[-1] != "]":
                    raise ParserError(f"line {no}: bad section")
                section = line[1:-1].strip()
                self.m_ini.setdefault(section, {})
                continue
            if "=" not in line:
                raise ParserError(f"line {no}: expected key = value")
            k, v = line.split("=", 1)
            self.m_ini.setdefault(section, {})[k.strip()] = v.strip()

    @classmethod
    def from_string(cls, text):
        return cls(text=text)

    def Has(self, section, prop):
        return prop in self.m_ini.get(section, {})

    def Get(self, section, prop):
        if section not in self.m_ini:
            raise ParserError(f"no section [{section}]")
        if prop not in self.m_ini[section]:
            raise ParserError(f"no key {prop} in [{section}]")
        v = self.m_ini[section][prop]
        try:
            return float(v)
        except ValueError as e:
            raise ParserError(f"[{section}] {prop} = '{v}' is not a number") from e


class Scales(enum.Enum):
    height = 0
    length = 1
    velocity = 2
    source = 3
    time = 4


class DimensionManager:
    """Scale(kind, x): solver units -> dimensional; Unscale: the inverse."""

    def __init__(self, parser):
        self.m_h0 = parser.Get("Dimensions", "h0")
        self.m_l0 = parser.Get("Dimensions", "l0")
        self.m_g = parser.Get("Dimensions", "g") if parser.Has("Dimensions", "g") else 1.0
        if not (self.m_h0 > 0 and self.m_l0 > 0 and self.m_g > 0):
            raise DimensionError("h0, l0 and g must be positive")
        self.m_c0 = math.sqrt(self.m_g * self.m_h0)

    def Scale(self, kind, x):
        h0, l0, c0 = self.m_h0, self.m_l0, self.m_c0
        return {Scales.height: x * h0, Scales.length: x * l0, Scales.velocity: x * c0,
                Scales.source: x * c0 / l0, Scales.time: x * l0 / c0}[kind]

    def Unscale(self, kind, x):
        h0, l0, c0 = self.m_h0, self.m_l0, self.m_c0
        return {Scales.height: x / h0, Scales.length: x / l0, Scales.velocity: x / c0,
                Scales.source: x * l0 / c0, Scales.time: x * c0 / l0}[kind]


def make_test(name, par, dim, mid=(0.0, 0.0)):
    """The analytic test `name` (LakeAtRest, Ball, ConservativeThacker,
    SolenoidalThacker, ClassicThacker) with its parameters read and unscaled
    as the reference's constructors do: cor, tau from [Common] as sources,
    delta as is; Ball x0, y0 (lengths, made relative to mid), u0, v0
    (velocities); Thacker H0 (height), p0, q0 (sources).  mid is in solver
    units."""
    U = dim.Unscale
    common = dict(mid=mid, cor=U(Scales.source, par.Get("Common", "cor")),
                  tau=U(Scales.source, par.Get("Common", "tau")))
    if name == "LakeAtRest":
        return analytic.LakeAtRest(**common)
    common["delta"] = par.Get("Common", "delta")
    if name == "Ball":
        return analytic.Ball(x0=U(Scales.length, par.Get("Ball", "x0")) - mid[0],
                             y0=U(Scales.length, par.Get("Ball", "y0")) - mid[1],
                             u0=U(Scales.velocity, par.Get("Ball", "u0")),
                             v0=U(Scales.velocity, par.Get("Ball", "v0")), **common)
    th = dict(H0=U(Scales.height, par.Get("Thacker", "H0")), p0=U(Scales.source, par.Get("Thacker", "p0")),
              q0=U(Scales.source, par.Get("Thacker", "q0")), **common)
    cls = {"ConservativeThacker": analytic.ConservativeThacker,
           "SolenoidalThacker": analytic.SolenoidalThacker,
           "ClassicThacker": analytic.ClassicThacker}.get(name)
    if cls is None:
        raise ValueError(f"unknown test {name}")
    return cls(**th)
