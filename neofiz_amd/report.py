"""Run output in the reference's formats: report/Csv.java, report/RunLog.java, report/Runs.java.

  Runs    the run directory, `runs` unless NEOFIZ_RUNS names another (the reference reads the
          system property neofiz.runs), created on first use.
  Csv     a table of doubles with optional row labels, every value at 17 significant digits
          so that it reads back bit for bit.
  RunLog  one line per run appended to runs/log.csv: when, program, key=value parameters and
          results; queried by program and by the whole parameter set.

`java_g` is java.util.Formatter's %.<p>g, which both Csv and RunLog (and the film's numbers)
print through. Java rounds the shortest decimal digits of the double half-up to p significant
figures and pads with zeros; C's printf rounds the exact binary value half-even instead, so
the two differ on ties and in the padding (Java writes 0.1 at 17 digits as
0.10000000000000000, C as 0.10000000000000001). Both read back to the same double.
"""
import math
import os
from dataclasses import dataclass, field
from datetime import datetime, timezone
from typing import Dict, List, Optional

from .solver import IllegalArgumentException


def _shortest(value):
    """(digits, point) of |value|'s shortest repr: value = 0.d1d2... x 10^(point + 1)."""
    text = repr(abs(float(value)))
    mant, _, exp = text.partition("e")
    e10 = int(exp) if exp else 0
    whole, _, frac = mant.partition(".")
    digits = (whole + frac).lstrip("0").rstrip("0") or "0"
    if whole.strip("0"):
        lead = len(whole.lstrip("0")) - 1
    else:
        lead = -(len(frac) - len(frac.lstrip("0")) + 1)
    return digits, lead + e10


def java_g(value, precision=6):
    """String.format(Locale.ROOT, "%.<precision>g", value) for a finite double."""
    p = max(1, precision)
    neg = math.copysign(1.0, value) < 0.0
    if value == 0.0:
        digits, point = "0", 0
    else:
        digits, point = _shortest(value)
        if len(digits) > p:  # half-up on the digit string
            head = int(digits[:p]) + (1 if digits[p] >= "5" else 0)
            if len(str(head)) > p:
                head //= 10
                point += 1
            digits = str(head)
    digits = digits.ljust(p, "0")
    if value != 0.0 and not (-4 <= point < p):
        out = digits[0] + ("." + digits[1:] if p > 1 else "") + "e" + ("-" if point < 0 else "+") + f"{abs(point):02d}"
    elif point >= 0:
        out = digits[:point + 1] + ("." + digits[point + 1:] if p > point + 1 else "")
    else:
        out = "0." + "0" * (-point - 1) + digits
    return "-" + out if neg else out


class Runs:
    """report/Runs.java:25-48."""
    PROPERTY = "NEOFIZ_RUNS"
    DEFAULT = "runs"

    @staticmethod
    def dir():
        d = os.environ.get(Runs.PROPERTY, Runs.DEFAULT)
        os.makedirs(d, exist_ok=True)
        return d

    @staticmethod
    def file(name):
        return os.path.join(Runs.dir(), name)


class Csv:
    """report/Csv.java:26-133."""

    def __init__(self, *headers):
        if not headers:
            raise IllegalArgumentException("a table needs columns")
        self.headers = list(headers)
        self.rows: List[List[float]] = []
        self.labels: List[Optional[str]] = []
        self.labelled = False

    @staticmethod
    def of(*headers):
        return Csv(*headers)

    def row(self, *values):
        """row(double...) or row(label, double...)."""
        if values and isinstance(values[0], str):
            label, vals = values[0], [float(v) for v in values[1:]]
            if len(vals) != len(self.headers) - 1:
                raise IllegalArgumentException(
                    f"row has {len(vals) + 1} values for {len(self.headers)} columns")
            self.labelled = True
        else:
            label, vals = None, [float(v) for v in values]
            if len(vals) != len(self.headers):
                raise IllegalArgumentException(f"row has {len(vals)} values for {len(self.headers)} columns")
        self.rows.append(vals)
        self.labels.append(label)
        return self

    def size(self):
        return len(self.rows)

    def write(self, name):
        return self.writeTo(Runs.file(name))

    def writeTo(self, path):
        parent = os.path.dirname(str(path))
        if parent:
            os.makedirs(parent, exist_ok=True)
        with open(path, "w", encoding="utf-8", newline="\n") as out:
            out.write(",".join(self.headers) + "\n")
            for vals, label in zip(self.rows, self.labels):
                cells = [Csv._quote(label or "")] if self.labelled else []
                cells += [Csv._format(v) for v in vals]
                out.write(",".join(cells) + "\n")
        return path

    @staticmethod
    def _quote(text):
        plain = not any(c in text for c in ',"\n\r')
        return text if plain else '"' + text.replace('"', '""') + '"'

    @staticmethod
    def _format(value):
        if math.isnan(value):
            return "nan"
        if math.isinf(value):
            return "inf" if value > 0.0 else "-inf"
        return java_g(value, 17)


@dataclass
class Entry:
    """RunLog.Entry (RunLog.java:60-78)."""
    when: str
    program: str
    parameters: Dict[str, str] = field(default_factory=dict)
    results: Dict[str, str] = field(default_factory=dict)

    def number(self, key):
        text = self.results.get(key)
        if text is None:
            return math.nan
        try:
            return float(text)
        except ValueError:
            return math.nan


class RunLog:
    """report/RunLog.java:34-211."""
    FILE = "log.csv"
    HEADERS = ("when", "program", "parameters", "results")

    def __init__(self, program):
        self.program = RunLog._check("program", program)
        self.parameters: Dict[str, str] = {}
        self.results: Dict[str, str] = {}

    @staticmethod
    def of(program):
        return RunLog(program)

    def parameter(self, key, value):
        text = RunLog._format(value) if not isinstance(value, str) else RunLog._check("value", value)
        self.parameters[RunLog._check("key", key)] = text
        return self

    def result(self, key, value):
        text = RunLog._format(value) if not isinstance(value, str) else RunLog._check("value", value)
        self.results[RunLog._check("key", key)] = text
        return self

    def commit(self):
        entry = Entry(_instant_now(), self.program, dict(self.parameters), dict(self.results))
        path = Runs.file(RunLog.FILE)
        fresh = not os.path.exists(path) or os.path.getsize(path) == 0
        line = (",".join(RunLog.HEADERS) + "\n") if fresh else ""
        line += f"{entry.when},{entry.program},{RunLog._pack(entry.parameters)},{RunLog._pack(entry.results)}\n"
        with open(path, "a", encoding="utf-8", newline="\n") as out:
            out.write(line)
        return entry

    @staticmethod
    def all():
        path = Runs.file(RunLog.FILE)
        if not os.path.exists(path):
            return []
        out = []
        with open(path, encoding="utf-8") as f:
            for line in f.read().splitlines():
                if not line.strip() or line.startswith(RunLog.HEADERS[0] + ","):
                    continue
                parts = line.split(",", 3)
                if len(parts) < 4:
                    continue
                out.append(Entry(parts[0], parts[1], RunLog._unpack(parts[2]), RunLog._unpack(parts[3])))
        return out

    @staticmethod
    def historyOf(program):
        return [e for e in RunLog.all() if e.program == program]

    def lastMatching(self):
        best = None
        for e in RunLog.all():
            if e.program == self.program and e.parameters == self.parameters:
                best = e
        return best

    @staticmethod
    def _pack(mapping):
        return ";".join(f"{k}={v}" for k, v in mapping.items())

    @staticmethod
    def _unpack(text):
        out = {}
        if not text:
            return out
        for pair in text.split(";"):
            at = pair.find("=")
            if at > 0:
                out[pair[:at]] = pair[at + 1:]
        return out

    @staticmethod
    def _format(value):  # :185-192
        value = float(value)
        if math.isnan(value):
            return "nan"
        if math.isinf(value):
            return "inf" if value > 0.0 else "-inf"
        if value == math.floor(value) and abs(value) < 1e15:
            return ("-" if math.copysign(1.0, value) < 0.0 else "") + str(int(abs(value)))  # %.0f
        return java_g(value, 6)

    @staticmethod
    def _check(what, text):  # :194-206
        if text is None or text == "":
            raise IllegalArgumentException(f"{what} must not be empty")
        if any(c in text for c in ",;=\n\r"):
            raise IllegalArgumentException(
                f"{what} must not contain a comma, semicolon, equals or newline; the log keeps "
                f"them as separators rather than escaping them: {text}")
        return text


def _instant_now():
    """Instant.now().toString(): ISO-8601 in UTC with a Z, the fraction in groups of three
    digits and dropped when zero."""
    t = datetime.now(timezone.utc)
    base = t.strftime("%Y-%m-%dT%H:%M:%S")
    us = t.microsecond
    if us == 0:
        return base + "Z"
    if us % 1000 == 0:
        return f"{base}.{us // 1000:03d}Z"
    return f"{base}.{us:06d}Z"
