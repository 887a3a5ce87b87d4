"""H.psi kernel families of one grid (sep_set_hpsi_family, sep_autotune_ex).

At complex64 128^2 the fused plane kernel (pass C + next pass A) has three
forms: the default on the separable LDS placement (k_plane_bwd_fwd_s), the
chunked one (k_plane_bwd_fwd_h: half a plane of LDS, two workgroups per CU)
and the xor-swizzled whole-plane one (k_plane_bwd_fwd_w).  All run the same
arithmetic up to the compiler's choice of fused multiply-adds, so a
complex64 filter at 128^3 (40 Newton steps, which amplify rounding) must lie
as close to the FP64 filter with each form as with the others, with one
lane and with two.  The default and FP64 families are checked against the
oracle at this size (test_gpu_fullsize.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx(sep, n, seed):
    rng = np.random.default_rng(seed)
    ctx = sep.Context(n, n, n, 0.7, 0.8, 0.9)
    ctx.set_potential(rng.uniform(-1.5, 1.0, n ** 3))
    return ctx, rng


def _filter(sep, ctx, psi0, prec, nc=40, ms=2):
    dE, _, _ = ctx.energy_range()
    zn, _ = sep.leja_points(nc)
    an = sep.newton_coeffs(zn, np.linspace(-0.5, 0.5, ms), dE, ctx.vmin, (nc / (2.5 * dE)) ** 2)
    out, _ = ctx.filter(psi0, an, zn, ctx.vmin, dE, prec)
    return out


def test_fused_kernel_forms_agree(sep):
    n = 128
    print("case fused kernel forms 128^3 complex64", flush=True)
    ctx, rng = _ctx(sep, n, 31)
    with ctx:
        psi0 = rng.standard_normal((8, n ** 3))
        psi0 /= np.linalg.norm(psi0, axis=1, keepdims=True)
        ref64 = _filter(sep, ctx, psi0, sep.SEP_PREC_FP64)
        scale = np.abs(ref64).max()
        got, err = {}, {}
        for fam in (sep.SEP_HPSI_PLANE, sep.SEP_HPSI_PLANE_SWIZZLED, sep.SEP_HPSI_PLANE_CHUNKED):
            ctx.set_hpsi_family(sep.SEP_PREC_FP32, fam)
            assert ctx.hpsi_family(sep.SEP_PREC_FP32) == fam
            for lanes in (1, 2):
                ctx.set_filter_lanes(lanes)
                out = _filter(sep, ctx, psi0, sep.SEP_PREC_FP32)
                got[fam, lanes] = out
                err[fam, lanes] = np.abs(out - ref64).max() / scale
                print(f"family {fam}, {lanes} lane(s): vs FP64 {err[fam, lanes]:.3e}", flush=True)
        worst = max(err.values())
        best = min(err.values())
        assert worst <= 1.5 * best + 1e-7
        ref = got[sep.SEP_HPSI_PLANE, 1]
        for key, out in got.items():
            d = np.abs(out - ref).max() / scale
            print(f"family {key[0]}, {key[1]} lane(s) vs the default form: {d:.3e}", flush=True)
            assert d < 1e-5


def test_family_availability(sep):
    """Plane families need a plane length; the chunked and swizzled forms are
    complex64 128^2 only."""
    with sep.Context(128, 128, 128, 0.7, 0.8, 0.9) as ctx:
        ctx.set_potential(np.zeros(128 ** 3))
        with pytest.raises(RuntimeError):
            ctx.set_hpsi_family(sep.SEP_PREC_FP64, sep.SEP_HPSI_PLANE)
        ctx.set_hpsi_family(sep.SEP_PREC_FP64, sep.SEP_HPSI_LINES)
        ctx.set_hpsi_family(sep.SEP_PREC_FP32, sep.SEP_HPSI_PLANE_CHUNKED)
        ctx.set_hpsi_family(sep.SEP_PREC_FP32, sep.SEP_HPSI_PLANE_SWIZZLED)
    with sep.Context(64, 64, 64, 0.7, 0.8, 0.9) as ctx:
        with pytest.raises(RuntimeError):
            ctx.set_hpsi_family(sep.SEP_PREC_FP32, sep.SEP_HPSI_PLANE_CHUNKED)
        with pytest.raises(RuntimeError):
            ctx.set_hpsi_family(sep.SEP_PREC_FP32, sep.SEP_HPSI_PLANE_SWIZZLED)
        ctx.set_hpsi_family(sep.SEP_PREC_FP32, sep.SEP_HPSI_PLANE)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_autotune_ex_picks_a_family_and_keeps_results(sep, prec):
    """The autotuner times every family the grid has with 1, 2 and 4 lanes,
    keeps the fastest, and the filter with its choice agrees with the
    default family's (complex64 1e-5 / FP64 1e-12)."""
    p = sep.SEP_PREC_FP32 if prec == "fp32" else sep.SEP_PREC_FP64
    n = 128 if prec == "fp32" else 96
    print("case autotune_ex", prec, n, flush=True)
    ctx, rng = _ctx(sep, n, 32)
    with ctx:
        psi0 = rng.standard_normal((8, n ** 3))
        psi0 /= np.linalg.norm(psi0, axis=1, keepdims=True)
        ref = _filter(sep, ctx, psi0, p)
        ms, fam, lanes = ctx.autotune_ex(p, nvec=4)
        print("autotune", prec, "ms per step per family", ms, "family", fam, "lanes", lanes,
              flush=True)
        assert fam in (0, 1, 2, 3) and ms[fam] > 0 and ms[fam] == ms[ms > 0].min()
        assert lanes in (1, 2, 4)
        assert ctx.hpsi_family(p) == fam
        got = _filter(sep, ctx, psi0, p)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print("after autotune: max rel vs default family", err, flush=True)
        # families differ by the compiler's choice of fused multiply-adds,
        # which 40 Newton steps amplify (the same bound as between the fused
        # kernel's forms above)
        assert err < (1e-5 if prec == "fp32" else 1e-12)
