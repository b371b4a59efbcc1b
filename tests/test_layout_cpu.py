"""Host check of the engine's buffer layout (fwi2d_amd/csrc/layout.h): every
zero-fill extent a pass or the autotuner issues lies inside the allocation
it writes, for sponge pads wider than, equal to and narrower than the
random-boundary pad, and for contexts with one grid only.

The header is compiled with g++ into a small driver that prints the sizes;
no GPU is involved.  The engine takes its allocation sizes and fill extents
from the same functions and checks each fill against the recorded size at
run time (FWI_ERR_INTERNAL on a violation)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include "layout.h"
#include <cstdio>
#include <cstdlib>
int main(int argc, char **argv)
{
    fwi_config c{};
    c.nz = atoi(argv[1]); c.nx = atoi(argv[2]);
    c.abc_nb = atoi(argv[3]); c.rbc_nb = atoi(argv[4]); c.max_batch = atoi(argv[5]);
    const fwi::Layout L = fwi::make_layout(c);
    printf("%d %ld %d %ld %ld %ld %ld %ld %ld %ld\n", L.grid[0].on, L.grid[0].fstride,
           L.grid[1].on, L.grid[1].fstride, L.fstride_max, L.wavefield(), L.rb_coef(),
           L.slot_acc(), L.abc_coef(), L.shot_sum());
    for (int n = 1; n <= c.max_batch; ++n)
        printf("%d %ld %ld %ld %ld\n", n, L.grid[0].on ? L.pass_fill(0, n) : 0L,
               L.grid[1].on ? L.pass_fill(1, n) : 0L, L.tune_fill(n),
               L.grid[1].on ? L.acc_fill(n) : 0L);
    return 0;
}
"""

GUARD = 1024


def fstride(nz, nx, nb):
    ldz = (nz + 2 * nb + 3) & ~3
    return (ldz * (nx + 2 * nb) + 63) & ~63


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("layout")
    src = d / "layout_driver.cpp"
    src.write_text(DRIVER)
    exe = d / "layout_driver"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "fwi2d_amd", "csrc"), str(src), "-o", str(exe)],
                   check=True)
    return str(exe)


def run(driver, nz, nx, abc_nb, rbc_nb, slots):
    out = subprocess.run([driver, str(nz), str(nx), str(abc_nb), str(rbc_nb), str(slots)],
                         check=True, capture_output=True, text=True).stdout.split("\n")
    head = [int(v) for v in out[0].split()]
    rows = [[int(v) for v in line.split()] for line in out[1:] if line.strip()]
    keys = ("on0", "fs0", "on1", "fs1", "fs_max", "wavefield", "rb_coef", "slot_acc",
            "abc_coef", "shot_sum")
    return dict(zip(keys, head)), rows


CASES = [
    (81, 201, 51, 41, 4),     # the reference's parameters.txt: abc 1250 m, rbc 1000 m at 25 m
    (81, 201, 41, 51, 4),
    (81, 201, 30, 30, 3),
    (301, 1001, 100, 150, 64),
    (401, 1601, 100, 200, 256),
    (256, 768, 120, 100, 96),
    (60, 120, 30, 0, 2),      # sponge grid only
    (60, 120, 0, 30, 2),      # random-boundary grid only
]


@pytest.mark.parametrize("nz,nx,abc_nb,rbc_nb,slots", CASES)
def test_fill_extents_stay_inside_their_allocations(driver, nz, nx, abc_nb, rbc_nb, slots):
    L, rows = run(driver, nz, nx, abc_nb, rbc_nb, slots)
    fs0 = fstride(nz, nx, abc_nb) if abc_nb else 0
    fs1 = fstride(nz, nx, rbc_nb) if rbc_nb else 0
    assert (L["fs0"] if L["on0"] else 0) == fs0 and (L["fs1"] if L["on1"] else 0) == fs1
    assert L["fs_max"] == max(fs0, fs1)
    assert L["wavefield"] == slots * max(fs0, fs1) + GUARD
    if rbc_nb:
        assert L["rb_coef"] == L["slot_acc"] == slots * fs1 + GUARD
        assert L["shot_sum"] == fs1
    if abc_nb:
        assert L["abc_coef"] == fs0 + GUARD
    assert len(rows) == slots
    for n, pass0, pass1, tune, acc in rows:
        # forward_batch / reverse_batch zero four fields of their grid
        assert pass0 <= L["wavefield"] and pass1 <= L["wavefield"]
        # the autotuner zeroes all eight fields for both grids' launches
        assert tune <= L["wavefield"] and tune >= max(pass0, pass1)
        if rbc_nb:
            # ... and the per-slot image / illumination of the rbc grid
            assert acc == n * fs1 and acc <= L["slot_acc"]


def test_the_wider_sponge_grid_would_overrun_the_accumulators(driver):
    """With abc_nb > rbc_nb the wavefield extent of a full batch exceeds the
    per-slot accumulators: a fill of the accumulators with that extent (what
    the autotuner once issued) would write past them."""
    L, rows = run(driver, 81, 201, 51, 41, 4)
    n, _, _, tune, acc = rows[-1]
    assert tune > L["slot_acc"] and acc <= L["slot_acc"]
