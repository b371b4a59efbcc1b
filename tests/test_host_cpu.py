"""CPU tests of the drop-in executables' host layer: the parameters.txt
parser, binary / text I/O and the failure paths that fire before any GPU
work (missing key, missing or short file, malformed geometry)."""
import os
import subprocess
import textwrap

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "fwi2d_amd", "host")
BIN = os.path.join(ROOT, "fwi2d_amd", "bin")

DRIVER = r"""
#include "admin.hpp"
#include <iostream>
int main(int argc, char** argv) {
    try {
        if (std::string(argv[1]) == "param") {
            std::cout << "[" << fwi::catch_parameter(argv[2], argv[3]) << "]" << std::endl;
        } else if (std::string(argv[1]) == "geom") {
            fwi::Geometry g; g.set_parameters(argv[2]);
            std::cout << g.nsrc << " " << g.nrec << " " << g.xsrc.back() << " "
                      << g.zrec.front() << std::endl;
        } else if (std::string(argv[1]) == "us") {    // survey.cpp seismogram file names
            const float dt = fwi::param_float("time_spacing", argv[2]);
            std::cout << std::to_string(int(1e6f * dt)) << std::endl;
        } else if (std::string(argv[1]) == "bin") {
            std::vector<float> a(std::stol(argv[3]));
            fwi::import_binary_float(argv[2], a.data(), (long)a.size());
            std::cout << a.back() << std::endl;
        }
    } catch (const std::exception& e) { std::cerr << e.what() << std::endl; return 1; }
    return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("admin")
    src = d / "drv.cpp"
    src.write_text(DRIVER)
    exe = d / "drv"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{HOST}", str(src),
                    os.path.join(HOST, "admin.cpp"), "-o", str(exe)], check=True)
    return str(exe)


def run(*args):
    return subprocess.run(list(args), capture_output=True, text=True)


PARAMS = textwrap.dedent("""\
    #------------------------------------------------------------
    # Model parameters
    #------------------------------------------------------------

    x_samples = 201
    z_samples = 81
    x_samples_extra = 7

    model_spacing = 25.0                # [m]
    model_file = ../inputs/models/overthrust_init_81x201_25m.bin
    quoted = "abc def"
    listed = [1, 2, 3]
    tight=5
       indented = 9
    #commented = 4
    empty =
    """)


def test_parameter_parser(driver, tmp_path):
    p = tmp_path / "parameters.txt"
    p.write_text(PARAMS)
    get = lambda k: run(driver, "param", k, str(p))
    assert get("x_samples").stdout.strip() == "[201]"
    assert get("x_samples_extra").stdout.strip() == "[7]"
    assert get("model_spacing").stdout.strip() == "[25.0]"
    assert get("model_file").stdout.strip() == "[../inputs/models/overthrust_init_81x201_25m.bin]"
    assert get("quoted").stdout.strip() == "[abcdef]"
    assert get("listed").stdout.strip() == "[1,2,3]"
    assert get("tight").stdout.strip() == "[5]"
    for missing in ("indented", "commented", "empty", "nope", "x_sample"):
        r = get(missing)
        assert r.returncode == 1 and "missing" in r.stderr, missing


def test_geometry_and_binary_io(driver, tmp_path):
    sps = tmp_path / "sps.txt"
    rps = tmp_path / "rps.txt"
    sps.write_text("# x,z\n1000.00,100.00\n\n4000.00,100.00\n")
    rps.write_text("0.00,0.00\n2500.00,0.00\n5000.00,12.50\n")
    p = tmp_path / "p.txt"
    p.write_text(f"SPS = {sps}\nRPS = {rps}\n")
    r = run(driver, "geom", str(p))
    assert r.returncode == 0 and r.stdout.split() == ["2", "3", "4000", "0"]
    b = tmp_path / "m.bin"
    np.arange(10, dtype=np.float32).tofile(b)
    assert run(driver, "bin", str(b), "10").stdout.strip() == "9"
    r = run(driver, "bin", str(b), "11")
    assert r.returncode == 1 and "fewer than 11" in r.stderr
    r = run(driver, "bin", str(tmp_path / "none.bin"), "1")
    assert r.returncode == 1 and "could not be opened" in r.stderr


@pytest.mark.parametrize("exe", ["modeling", "migration", "inversion"])
def test_mains_fail_cleanly_before_gpu_work(exe, tmp_path):
    path = os.path.join(BIN, exe)
    assert os.path.exists(path), "build the executables first"
    r = run(path)
    assert r.returncode == 2 and "usage" in r.stderr
    p = tmp_path / "p.txt"
    p.write_text("x_samples = 20\n")          # everything else missing
    r = run(path, str(p))
    assert r.returncode == 1 and "missing" in r.stderr
    r = run(path, str(tmp_path / "absent.txt"))
    assert r.returncode == 1 and "could not be opened" in r.stderr


def test_viewer_parser_agrees_with_the_executables(driver, tmp_path):
    """fwi2d_amd.viewer reads parameters.txt with the executables' rules."""
    from fwi2d_amd import viewer
    p = tmp_path / "parameters.txt"
    p.write_text(PARAMS)
    for key in ("x_samples", "x_samples_extra", "z_samples", "model_spacing", "model_file",
                "quoted", "listed", "tight", "indented", "commented", "empty", "nope",
                "x_sample"):
        r = run(driver, "param", key, str(p))
        if r.returncode == 0:
            assert "[" + viewer.read_parameter(str(p), key) + "]" == r.stdout.strip(), key
        else:
            with pytest.raises(KeyError):
                viewer.read_parameter(str(p), key)
            assert viewer.read_parameter(str(p), key, "") == "", key


@pytest.mark.parametrize("dt", ["1e-3", "8e-4", "6e-4", "5e-4", "3e-4", "0.00025", "7e-5"])
def test_prep_file_names_agree_with_the_executables(driver, tmp_path, dt):
    """fwi2d_amd.configs names seismogram files with the executables' float32
    rounding of 1e6 * dt."""
    from fwi2d_amd import configs as C
    p = tmp_path / "p.txt"
    p.write_text(f"time_spacing = {dt}\n")
    us = run(driver, "us", str(p)).stdout.strip()
    assert C.seismogram_prefix(2001, 51, float(dt)) == f"seismogram_nt2001_nr51_{us}us_shot_"
