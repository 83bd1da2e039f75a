"""sph_metrics_noalloc (the allocation-free entry point) against sph_metrics,
through the C ABI: two calls in a row, the second on a smaller scene, must give
the same bits as sph_metrics on the same state; a scene beyond the stated
limit is refused with SPH_ERR_INVALID."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("mean_rho", "kinetic", "max_speed", "nonfinite")


def _same(a, b):
    for f in FIELDS:
        assert np.float64(getattr(a, f)).tobytes() == np.float64(getattr(b, f)).tobytes(), f
    assert np.array(a.com[:]).tobytes() == np.array(b.com[:]).tobytes()


def test_noalloc_metrics_equal_sph_metrics(pkg, gpu_lib):
    sph = pkg.sph
    big, small = pkg.scenes.dam1m(), pkg.scenes.dam10k()
    counts = [big.count, small.count]
    bounds = [sph.make_boundary(s.lower, s.upper) for s in (big, small)]
    with pkg.Simulation(counts, boundaries=bounds) as sim:
        sim.init_lattice(sph.make_lattice(**big.lattice_args()), scene=0)
        sim.init_lattice(sph.make_lattice(**small.lattice_args()), scene=1)
        sim.step(300)  # collapsed, with the reference's coincident-pair NaNs in scene 0
        for _ in range(2):
            a0 = sim.metrics_noalloc(0)
            a1 = sim.metrics_noalloc(1)  # second call, smaller input
            _same(a0, sim.metrics(0))
            _same(a1, sim.metrics(1))
            assert a0.mean_rho > 1000.0 and a1.mean_rho > 1000.0
            sim.step(3)
        # against numpy on the downloaded state (float64 sums, loose tolerance)
        st = sim.download()[counts[0]:]
        m = sim.metrics_noalloc(1)
        assert m.mean_rho == pytest.approx(st["extras"][:, 0].astype(np.float64).mean(), rel=1e-12)


def test_noalloc_metrics_refuses_beyond_limit(pkg, gpu_lib):
    limit = 16384 * 256
    with pkg.Simulation([limit + 1, 10]) as sim:
        with pytest.raises(pkg.SphError) as e:
            sim.metrics_noalloc(0)
        assert e.value.code == pkg.sph.SPH_ERR_INVALID and "4194304" in str(e.value)
        m = sim.metrics_noalloc(1)  # the context stays usable
        _same(m, sim.metrics(1))
