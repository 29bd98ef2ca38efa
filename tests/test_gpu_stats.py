"""Device statistics sums (pll_stat_*): the calcStat quantities of
src/mod_statistics.cc:30-106 sampled on the device, checked bit for bit
against numpy on the downloaded fields."""
import numpy as np
import pytest

from cases import make_channel, make_oracle

pytestmark = pytest.mark.gpu


def _quantities(u, v, w, p, u_crf=0.0):
    uc = (u[1:-1, 1:-1, :-2] + u[1:-1, 1:-1, 1:-1]) * 0.5 + u_crf
    vc = (v[1:-1, :-2, 1:-1] + v[1:-1, 1:-1, 1:-1]) * 0.5
    wc = (w[:-2, 1:-1, 1:-1] + w[1:-1, 1:-1, 1:-1]) * 0.5
    pc = p[1:-1, 1:-1, 1:-1]
    return (uc, vc, wc, uc * uc, vc * vc, wc * wc, uc * vc, uc * wc, vc * wc, pc, pc * pc)


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
@pytest.mark.parametrize("name", ["r235", "gen"])
def test_stat_sums_bitexact(name, precision):
    print(f"[stats] {name} {precision}", flush=True)
    o = make_oracle(name)
    u, v, w = o.field(), o.field(), o.field()
    o.init_flow(u, v, w, "poi", u0=1.0, with_noise=True, noise_intensity=0.1)
    with make_channel(name, precision=precision) as ch:
        for f, a in zip("uvw", (u, v, w)):
            ch.upload(f, a)
        ch.updateBoundVel(0)
        which = [1] * 11
        which[4] = 0  # an unselected quantity stays unallocated
        ch.stat_enable(which)
        sums = [np.zeros((ch.nz, ch.ny, ch.nx)) for _ in range(11)]
        for _ in range(3):
            ch.step(1)
            fields = [ch.download(f) for f in ("u", "v", "w", "p")]
            for t, q in enumerate(_quantities(*fields)):
                sums[t] += q
            ch.stat_accumulate()
        for t in range(11):
            if not which[t]:
                with pytest.raises(Exception):
                    ch.stat_download(t)
                continue
            np.testing.assert_array_equal(ch.stat_download(t), sums[t], err_msg=str(t))
            np.testing.assert_allclose(ch.stat_plane_sums(t), sums[t].sum(axis=(1, 2)),
                                       rtol=1e-12, atol=1e-12 * np.abs(sums[t]).max())
        # upload round trip, then re-enable zeroes
        ch.stat_upload(0, sums[3])
        np.testing.assert_array_equal(ch.stat_download(0), sums[3])
        ch.stat_enable(which)
        assert not ch.stat_download(0).any()
