"""The batched chemistry kernel's block split (mvd_chem_dirk_fit_block), on the CPU: the
launch asks the runtime for the kernel's static LDS and the device's per-block limit and
halves the block until the dynamic LDS (gamma tables, then per wave its cells' constants,
earlier stage derivatives and step start) plus the static part fits. Tables of more than
four stages keep their stage derivatives in a global buffer instead of LDS."""
import pytest

from conftest import load_pyapi

mvd = load_pyapi()
GFX950_LDS = 160 * 1024
STATIC = 20 * 4                     # the kernel's block statistics (20 ints)


def dyn_bytes(block, nstage):
    waves = block // 64
    fst = 10 * (nstage - 1) if nstage <= 4 else 0
    return 8 * (4 * 1024 + waves * (24 + fst + 10) * 64)


@pytest.mark.parametrize("nstage", [1, 2, 3, 4, 5, 7])
@pytest.mark.parametrize("block", [64, 128, 256])
@pytest.mark.parametrize("limit", [GFX950_LDS, 64 * 1024, 48 * 1024])
def test_fit_block_is_the_largest_that_fits(nstage, block, limit):
    got = mvd.lib().mvd_chem_dirk_fit_block(block, nstage, STATIC, limit)
    if got == 0:
        assert dyn_bytes(64, nstage) + STATIC > limit
        return
    assert got in (64, 128, 256) and got <= block
    assert dyn_bytes(got, nstage) + STATIC <= limit
    if got < block:
        assert dyn_bytes(2 * got, nstage) + STATIC > limit


def test_fit_block_known_cases():
    f = mvd.lib().mvd_chem_dirk_fit_block
    # three stages: four waves take 140 KiB of LDS, a 256-thread block fits a gfx950 CU
    assert f(256, 3, STATIC, GFX950_LDS) == 256
    # four stages: four waves take exactly 160 KiB of dynamic LDS, so with the static block
    # statistics (80 bytes) the block must split into two waves
    assert dyn_bytes(256, 4) == GFX950_LDS
    assert f(256, 4, STATIC, GFX950_LDS) == 128
    assert f(256, 4, 0, GFX950_LDS) == 256
    # a 64 KiB limit leaves one wave per block; 32 KiB cannot hold the gamma tables and one
    # wave
    assert f(256, 3, STATIC, 64 * 1024) == 64
    assert f(256, 1, STATIC, 32 * 1024) == 0
    # seven stages (ESDIRK ARK4(3)7L[2]SA): the stage derivatives are global, four waves fit
    assert f(256, 7, STATIC, GFX950_LDS) == 256


def test_fit_block_rejects_bad_arguments():
    f = mvd.lib().mvd_chem_dirk_fit_block
    assert f(96, 3, 0, GFX950_LDS) == mvd.MVD_BAD_ARGUMENT
    assert f(256, 8, 0, GFX950_LDS) == mvd.MVD_BAD_ARGUMENT
    assert f(256, 0, 0, GFX950_LDS) == mvd.MVD_BAD_ARGUMENT
    assert f(256, 3, -1, GFX950_LDS) == mvd.MVD_BAD_ARGUMENT
