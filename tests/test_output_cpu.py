"""Output modules on the host (test/testOutput.cpp:37-256): the reference's header
tests for each output type, properties, the version line, file errors, and the
ParticleCollector; plus the number formats of every column (src/module/TextOutput.cpp)."""
import os

import pytest

from crpropa3_amd import (Candidate, EeV, Mpc, Myr, Output, ParticleCollector, TextOutput, Vector3d, nucleusId)


def first_line(capfd, output):
    c = Candidate()
    output.process(c)
    output.close()  # flushes standard output
    out = capfd.readouterr().out
    return out, out[:out.find("\n")]


def test_output_size():
    c = Candidate()
    o = Output()
    for _ in range(5):
        o.process(c)
    assert o.size() == 5


@pytest.mark.parametrize("otype,header", [
    (Output.Trajectory1D, "#\tID\tE\tX"),
    (Output.Event1D, "#\tD\tID\tE\tID0\tE0"),
    (Output.Trajectory3D, "#\tD\tID\tE\tX\tY\tZ\tPx\tPy\tPz"),
    (Output.Event3D, "#\tD\tID\tE\tX\tY\tZ\tPx\tPy\tPz\tID0\tE0\tX0\tY0\tZ0\tP0x\tP0y\tP0z")])
def test_print_header(capfd, otype, header):
    assert first_line(capfd, TextOutput(otype))[1] == header


def test_print_header_custom(capfd):
    o = TextOutput(Output.Event1D)
    o.enable(Output.SerialNumberColumn)
    o.disable(Output.TrajectoryLengthColumn)
    o.set(Output.RedshiftColumn, False)
    o.enable(Output.CandidateTagColumn)
    o.enable(Output.TimeColumn)
    assert first_line(capfd, o)[1] == "#\ttime\tSN\tID\tE\tSN0\tID0\tE0\tSN1\ttag"


def test_print_property(capfd):
    o = TextOutput(Output.Event1D)
    o.disableAll()
    o.enableProperty("foo", 2.0, "Bar")
    assert first_line(capfd, o)[1] == "#\tfoo"


def test_print_header_version(capfd):
    out, _ = first_line(capfd, TextOutput(Output.Event1D))
    i = out.find("# CRPropa version: ") + 19
    assert out[i:out.find("\n", i)].startswith("crpropa3_amd")


def test_fail_on_illegal_output_file():
    with pytest.raises(RuntimeError):
        TextOutput("THIS_FOLDER_MUST_NOT_EXISTS_12345+/FILE.txt")


def test_configuration_fixed_after_writing(capfd):
    o = TextOutput(Output.Event1D)
    o.process(Candidate())
    with pytest.raises(RuntimeError, match="after data has been written"):
        o.enable(Output.TimeColumn)


def test_trajectory3d_line_formats(capfd):
    """%8.5E for lengths, energies, positions and directions, %10i for ids, tab separated."""
    c = Candidate(nucleusId(4, 2), 12.3456789 * EeV, Vector3d(1.5, -2.25, 1e-3) * Mpc, Vector3d(0, 0.6, 0.8))
    c.setTrajectoryLength(123.456789 * Mpc)
    o = TextOutput(Output.Trajectory3D)
    o.process(c)
    o.close()
    line = [ln for ln in capfd.readouterr().out.splitlines() if not ln.startswith("#")][0]
    d = c.current.getDirection()
    expect = "\t".join(["%8.5E" % 123.456789, "%10i" % nucleusId(4, 2), "%8.5E" % 12.3456789,
                        "%8.5E" % 1.5, "%8.5E" % -2.25, "%8.5E" % 1e-3, "%8.5E" % d.x, "%8.5E" % d.y, "%8.5E" % d.z])
    assert line == expect


def test_particle_collector_size_fetch_reprocess():
    c = Candidate(nucleusId(1, 1), 1 * EeV)
    out = ParticleCollector()
    for _ in range(5):
        out.process(c)
    assert out.size() == 5 and len(out) == 5
    assert out[0].getSerialNumber() == c.getSerialNumber()
    again = ParticleCollector()
    out.reprocess(again)
    assert again.size() == 5 and again[0].getSerialNumber() == c.getSerialNumber()


def test_particle_collector_dump_load(tmp_path):
    c = Candidate(nucleusId(1, 1), 1.234 * EeV)
    c.current.setPosition(Vector3d(1, 2, 3))
    c.current.setDirection(Vector3d(-1, -1, -1))
    c.setTrajectoryLength(1 * Mpc)
    c.setRedshift(2)
    c.setTime(3 * Myr)
    inp, out = ParticleCollector(), ParticleCollector()
    for _ in range(11):
        inp.process(c)
    fn = str(tmp_path / "ParticleCollector_DumpTest.txt")
    inp.dump(fn)
    out.load(fn)
    assert inp.size() == out.size()
    assert out[0].current.getEnergy() == pytest.approx(c.current.getEnergy(), rel=1e-5)
    assert out[1].getTrajectoryLength() == pytest.approx(c.getTrajectoryLength(), rel=1e-5)
    assert out[2].current.getId() == c.current.getId()
    assert out[3].getRedshift() == c.getRedshift()
    assert out[4].getTime() == pytest.approx(c.getTime(), rel=1e-5)
    assert os.path.getsize(fn) > 0
