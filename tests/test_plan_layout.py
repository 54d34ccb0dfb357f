"""CPU tests of the host-side plan layout (pvd_plan_describe: no device call).

They pin the choices the dose kernels rely on: FFT lengths that never alias, the
transform path of each axis, and the guard that keeps the register-resident
256-point X pass off volumes whose byte offsets would not fit in 32 bits.
"""
import itertools

import numpy as np
import pytest

from pyvoxeldosimetry_amd import describe_plan


def smooth(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def test_ct_and_torso_keep_their_register_paths():
    d = describe_plan((256, 256, 200), (31, 31, 31), "reference", algo="fft")
    assert d["fft_n"] == (256, 256, 200)
    assert d["fft_path"] == ("sq256", "sq256", "two_step")
    d = describe_plan((512, 512, 400), (61, 61, 61), "linear", algo="fft")
    assert d["fft_n"] == (560, 560, 432)
    assert d["fft_path"] == ("two_step", "two_step", "two_step")


def x_offsets_fit(nx, px, py, hp):
    """The 32-bit byte offsets of the register X passes: the x-plane walk and the
    general kernel spectrum, plus a tile's columns (restated from the kernels'
    addressing)."""
    wplane = py * hp
    lim = 2 ** 32 - 1 - 256 - 8 * hp
    return (nx - 1) * wplane * 8 <= lim and (px - 1) * py * hp * 8 <= lim


@pytest.mark.parametrize("shape", [(256, 2048, 2048), (256, 2048, 1800), (256, 1600, 2048), (256, 2048, 1024),
                                   (256, 256, 200), (256, 1024, 2000)])
def test_256_point_x_pass_only_where_32_bit_offsets_fit(shape):
    d = describe_plan(shape, (3, 3, 3), "reference", algo="fft")
    px, py, pz = d["fft_n"]
    assert px == 256
    h = pz // 2 + 1
    hp = -(-h // 16) * 16
    fits = x_offsets_fit(shape[0], px, py, hp)
    assert (d["fft_path"][0] == "sq256") == fits, (shape, d["fft_path"], fits)
    if shape == (256, 2048, 2048):
        assert not fits                    # 255 * 2048 * 1040 * 8 bytes > 4 GiB


XY_TWO_STEP_B = [6, 8, 9, 10, 12, 14, 15, 16, 18, 20, 21, 24, 25, 27, 28, 30, 32, 35, 36]


@pytest.mark.parametrize("b", XY_TWO_STEP_B)
def test_plan_picks_8_column_x_tiles_for_the_fused_two_step_from_288_to_512_points(b):
    """An x tile left to the plan is 8 columns for the fused two-step lengths 16 * B with
    18 <= B <= 32 (the LDS-resident 560/576 and every shorter length keep 16); the y tile,
    an explicit x tile and the Stockham flag are left alone."""
    P = 16 * b
    d = describe_plan((P, 6, 10), (5, 3, 3), "reference", algo="fft")
    assert d["fft_n"][0] == P
    assert d["fft_path"][0] == ("sq256" if P == 256 else "two_step")
    assert d["tile"][0] == (8 if 18 <= b <= 32 else 16), d
    assert describe_plan((6, P, 10), (3, 5, 3), "reference", algo="fft")["tile"][1] == 16
    assert describe_plan((P, 6, 10), (5, 3, 3), "reference", algo="fft", tile=(16, 16, 16))["tile"][0] == 16
    assert describe_plan((P, 6, 10), (5, 3, 3), "reference", algo="fft", stockham=True)["tile"][0] == 16


@pytest.mark.parametrize("shape", [(384, 2048, 2048), (384, 1024, 1024), (320, 1600, 1800), (448, 512, 512)])
def test_fused_two_step_x_pass_only_where_32_bit_offsets_fit(shape):
    d = describe_plan(shape, (3, 3, 3), "reference", algo="fft")
    px, py, pz = d["fft_n"]
    assert px == shape[0] and px % 16 == 0 and px // 16 < 35
    h = pz // 2 + 1
    hp = -(-h // 16) * 16
    fits = x_offsets_fit(shape[0], px, py, hp)
    assert (d["fft_path"][0] == "two_step") == fits, (shape, d["fft_path"], fits)
    if shape == (384, 2048, 2048):
        assert not fits                    # 383 * 2048 * 1040 * 8 bytes > 4 GiB


def test_register_padding_never_aliases_and_smooth_lengths_flag_restores_the_7_smooth_rule():
    rng = np.random.default_rng(7)
    for _ in range(60):
        shape = tuple(int(x) for x in rng.integers(20, 420, size=3))
        k = int(rng.choice([3, 11, 31, 61]))
        for mode in ("reference", "linear"):
            d = describe_plan(shape, (k, k, k), mode, algo="fft")
            s = describe_plan(shape, (k, k, k), mode, algo="fft", smooth_lengths=True)
            for a in range(3):
                n = shape[a]
                ku = min(k, n) if mode == "reference" else None
                # linear: outputs reach kernel taps within min(K//2, N-1) of the centre
                need = n + ku - 1 if mode == "reference" else n + min(k // 2, n - 1)
                p = d["fft_n"][a]
                assert smooth(p)
                if mode == "reference":
                    assert p == n or p >= need
                else:
                    assert p >= need
                p0 = (n if (mode == "reference" and smooth(n)) else next(m for m in itertools.count(need)
                                                                        if smooth(m)))
                assert s["fft_n"][a] == p0
                assert p <= 2 * p0
                if p != p0:
                    assert p0 > 128 and d["fft_path"][a] != "stockham"


def test_padding_reaches_register_paths_on_uneven_batch_shapes():
    from pyvoxeldosimetry_amd.phantoms import batch_specs
    n_reg = n_all = 0
    for spec in batch_specs():
        d = describe_plan(spec.shape, spec.kernel_grid, "reference", algo="fft")
        n_all += 3
        n_reg += sum(p != "stockham" for p in d["fft_path"])
    assert n_reg >= 0.75 * n_all, (n_reg, n_all)


def test_describe_rejects_bad_parameters():
    with pytest.raises(ValueError):
        describe_plan((0, 8, 8), (3, 3, 3))
    with pytest.raises(ValueError):
        describe_plan((8, 8, 8), (3, 3, 3), fft_n=(9, 8, 8))       # aliases: needs 8 or >= 10
    with pytest.raises(NotImplementedError):
        describe_plan((64, 64, 64), (31, 31, 31), algo="direct")    # 31 > 16 taps
