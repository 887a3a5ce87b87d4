"""Device memory after a filter: the ring of Newton vectors (sized from the
free memory and kept between filters) is released when a later allocation
of the context would not fit otherwise, and a new filter allocates it again."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_idle_ring_released_for_later_allocations(sep):
    import torch
    n, ns, ms, nc = 64, 16, 8, 1024
    print("case ring release 64^3", flush=True)
    rng = np.random.default_rng(5)
    with sep.Context(n, n, n, 0.7, 0.7, 0.7) as ctx:
        ctx.set_potential(rng.uniform(-1.0, 0.5, n ** 3))
        dE, _, _ = ctx.energy_range()
        zn, _ = sep.leja_points(nc)
        an = sep.newton_coeffs(zn, np.linspace(-0.5, 0.0, ms), dE, ctx.vmin,
                               (nc / (2.5 * dE)) ** 2)
        psi0 = rng.standard_normal((ns, n ** 3))
        ref, _ = ctx.filter(psi0, an, zn, ctx.vmin, dE, sep.SEP_PREC_FP64)
        ring = 1024 * ns * n ** 3 * 8  # K = 1024 FP64 Newton vectors per state
        free0 = torch.cuda.mem_get_info()[0]
        # all but 1 GB of the device taken: 1000 resident states (2 GB) fit
        # only where the idle ring goes
        block = torch.empty(free0 - (1 << 30), dtype=torch.uint8, device="cuda:0")
        big = rng.standard_normal((1000, n ** 3))
        ctx.states_upload(big)
        free1 = torch.cuda.mem_get_info()[0]
        print(f"ring {ring / 1e9:.1f} GB; free after the upload {free1 / 1e9:.1f} GB", flush=True)
        assert free1 > ring - (4 << 30)
        back = ctx.states_download(3)
        assert np.array_equal(back, big[:3])
        del block
        torch.cuda.empty_cache()
        # the next filter allocates its ring again and gives the same states
        got, _ = ctx.filter(psi0, an, zn, ctx.vmin, dE, sep.SEP_PREC_FP64)
        assert np.array_equal(got, ref)
