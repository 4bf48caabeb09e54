"""Run output: the reference's CsvTest and RunLogTest (src/test/java/org/neofiz/report/)
restated on neofiz_amd.report, and java_g against Java's %g conversion at the precisions the
reports use."""
import math
import os

import numpy as np
import pytest

from neofiz_amd.report import Csv, RunLog, Runs, java_g
from neofiz_amd.solver import IllegalArgumentException


@pytest.fixture
def runs(tmp_path, monkeypatch):
    monkeypatch.setenv(Runs.PROPERTY, str(tmp_path))
    return tmp_path


# ---------------------------------------------------------------- Csv (CsvTest.java)

def test_exact_round_trip(tmp_path):  # :29-52
    awkward = [0.1, 1.0 / 3.0, 18067432.198765431, 1e-300, 1e300, math.pi, np.nextafter(1.0, 2.0), -0.0,
               5e-324, 1.7976931348623157e308]
    csv = Csv.of("v")
    for v in awkward:
        csv.row(v)
    lines = open(csv.writeTo(str(tmp_path / "round-trip.csv"))).read().splitlines()
    assert len(lines) == len(awkward) + 1
    for v, line in zip(awkward, lines[1:]):
        assert np.float64(float(line)).view(np.uint64) == np.float64(v).view(np.uint64), (v, line)


def test_locale_independent_and_header(tmp_path):  # :53-82
    path = Csv.of("a", "b").row(1.5, 2.25).writeTo(str(tmp_path / "locale.csv"))
    row = open(path).read().splitlines()[1]
    assert row.count(",") == 1 and float(row.split(",")[0]) == 1.5 and float(row.split(",")[1]) == 2.25
    path = Csv.of("time_s", "capacity_Pa").row(0.0, 1.0).writeTo(str(tmp_path / "header.csv"))
    assert open(path).read().splitlines()[0] == "time_s,capacity_Pa"


def test_non_finite_values_are_named(tmp_path):  # :83-92
    path = Csv.of("v").row(math.nan).row(math.inf).row(-math.inf).writeTo(str(tmp_path / "edges.csv"))
    assert open(path).read().splitlines() == ["v", "nan", "inf", "-inf"]


def test_ragged_rows_are_refused():  # :94-103
    csv = Csv.of("a", "b")
    with pytest.raises(IllegalArgumentException, match="1 values for 2 columns"):
        csv.row(1.0)
    with pytest.raises(IllegalArgumentException):
        csv.row(1.0, 2.0, 3.0)
    with pytest.raises(IllegalArgumentException):
        Csv.of()


def test_parent_created_and_runs_directory(tmp_path, monkeypatch):  # :105-133
    path = Csv.of("x").row(1.0).writeTo(str(tmp_path / "deep" / "er" / "t.csv"))
    assert os.path.exists(path)
    monkeypatch.delenv(Runs.PROPERTY, raising=False)
    monkeypatch.chdir(tmp_path)
    assert os.path.basename(os.path.dirname(os.path.abspath(Runs.file("trace.csv")))) == "runs"
    monkeypatch.setenv(Runs.PROPERTY, str(tmp_path / "elsewhere"))
    assert Runs.file("trace.csv") == os.path.join(str(tmp_path / "elsewhere"), "trace.csv")
    assert os.path.isdir(tmp_path / "elsewhere")


def test_labelled_rows(tmp_path):  # :135-161
    path = (Csv.of("configuration", "rate").row("bare elements", 9.0).row("with, commas", 3.5)
            .writeTo(str(tmp_path / "labelled.csv")))
    lines = open(path).read().splitlines()
    assert lines[0] == "configuration,rate"
    assert lines[1].startswith("bare elements,9.0")
    assert lines[2].startswith('"with, commas",3.5')
    csv = Csv.of("a", "b", "c")
    with pytest.raises(IllegalArgumentException):
        csv.row("x", 1.0)
    with pytest.raises(IllegalArgumentException):
        csv.row("x", 1.0, 2.0, 3.0)


# ---------------------------------------------------------------- RunLog (RunLogTest.java)

def test_round_trip(runs):  # :50-64
    RunLog.of("smash").parameter("speed", 700.0).parameter("cell", 0.001) \
        .result("eroded", 32).result("plastic", 1618.9).commit()
    entries = RunLog.all()
    assert len(entries) == 1
    e = entries[0]
    assert e.program == "smash" and e.parameters["speed"] == "700" and e.parameters["cell"] == "0.00100000"
    assert e.number("eroded") == 32.0 and abs(e.number("plastic") - 1618.90) <= 0.01


def test_matching_is_on_the_whole_parameter_set(runs):  # :66-79
    RunLog.of("smash").parameter("speed", 700.0).result("eroded", 32).commit()
    assert RunLog.of("smash").parameter("speed", 700.0).lastMatching() is not None
    assert RunLog.of("smash").parameter("speed", 420.0).lastMatching() is None
    assert RunLog.of("topple").parameter("speed", 700.0).lastMatching() is None
    assert RunLog.of("smash").parameter("speed", 700.0).parameter("cell", 0.001).lastMatching() is None


def test_matching_returns_the_latest(runs):  # :81-90
    RunLog.of("smash").parameter("speed", 700.0).result("eroded", 32).commit()
    RunLog.of("smash").parameter("speed", 700.0).result("eroded", 41).commit()
    assert RunLog.of("smash").parameter("speed", 700.0).lastMatching().number("eroded") == 41.0


def test_history_is_per_program(runs):  # :92-103
    RunLog.of("smash").parameter("speed", 700.0).commit()
    RunLog.of("topple").parameter("blocks", 5).commit()
    RunLog.of("smash").parameter("speed", 420.0).commit()
    assert len(RunLog.all()) == 3
    assert len(RunLog.historyOf("smash")) == 2 and len(RunLog.historyOf("topple")) == 1
    assert RunLog.historyOf("nothing") == []


def test_the_file_is_readable(runs):  # :105-119
    RunLog.of("smash").parameter("speed", 700.0).result("eroded", 32).commit()
    RunLog.of("smash").parameter("speed", 420.0).result("eroded", 6).commit()
    lines = open(runs / RunLog.FILE).read().splitlines()
    assert len(lines) == 3 and lines[0] == "when,program,parameters,results"
    assert lines[1].endswith(",smash,speed=700,eroded=32") and lines[2].endswith(",smash,speed=420,eroded=6")
    when = lines[1].split(",")[0]
    assert when.endswith("Z") and when[10] == "T"


def test_separators_are_refused(runs):  # :121-133
    log = RunLog.of("smash")
    for key, value in (("a,b", "1"), ("a;b", "1"), ("a=b", "1"), ("ok", "a;b"), ("", "1")):
        with pytest.raises(IllegalArgumentException):
            log.parameter(key, value)
    with pytest.raises(IllegalArgumentException):
        RunLog.of("with,comma")


def test_empty_log_reads_as_empty(runs):  # :135-140
    assert RunLog.all() == []
    assert RunLog.of("smash").lastMatching() is None


# ---------------------------------------------------------------- Java's %g

@pytest.mark.parametrize("value,precision,text", [
    (0.001, 6, "0.00100000"), (1618.9, 6, "1618.90"), (0.1, 17, "0.10000000000000000"),
    (1.0 / 3.0, 17, "0.33333333333333330"), (1e-300, 17, "1.0000000000000000e-300"),
    (1e300, 17, "1.0000000000000000e+300"), (-0.0, 17, "-0.0000000000000000"), (0.0, 6, "0.00000"),
    (123456789.0, 6, "1.23457e+08"), (1e-5, 6, "1.00000e-05"), (999999.5, 6, "1.00000e+06"),
    (100000.5, 6, "100001"), (1.000005, 6, "1.00001"), (2.5e-7, 6, "2.50000e-07"), (0.0001, 6, "0.000100000"),
    (18067432.198765431, 17, "18067432.198765430"),
])
def test_java_general_format(value, precision, text):
    """Formatter's GENERAL conversion: decimal when 1e-4 <= |x| < 10^p after rounding, the
    shortest digits rounded half-up and padded with zeros, a signed two-digit exponent."""
    assert java_g(value, precision) == text
