"""The no-allocation direct-gravity entry point against the allocating one, through the C-ABI.

isph_gravity_direct_noalloc takes device arrays (allocated here by torch, as a caller's would be) and its
own fixed scratch; its output must match isph_gravity_accelerations(mode ISPH_GRAVITY_DIRECT) on the same
bodies to the f32 direct sum's tolerance, on two calls in a row (the second smaller, so stale scratch
contents past the new size would show), and reject a size beyond its scratch with ISPH_ERANGE.
"""
import ctypes

import numpy as np
import pytest

from botheadstudios_amd import isph
from oracle import oracle

from helpers import rel_err_floor, xorshift_debris

pytestmark = pytest.mark.gpu

SOFT = 2.5e4
# f32 sums of the same terms in a different order (Morton order in the allocating call, input order with
# compensated sums here): per-particle floored relative difference; both within 1e-5 of the f64 sum too
TOL = 1e-5


def _noalloc(lib, torch, pos, mass):
    n = len(mass)
    dp = torch.tensor(np.ascontiguousarray(pos, np.float64), device="cuda")
    dm = torch.tensor(np.ascontiguousarray(mass, np.float64), device="cuda")
    da = torch.full((n, 3), float("nan"), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    rc = lib.isph_gravity_direct_noalloc(dp.data_ptr(), dm.data_ptr(), n, SOFT, da.data_ptr(), stream.cuda_stream)
    assert rc == isph.OK, lib.isph_last_error(None).decode()
    stream.synchronize()
    return da.cpu().numpy()


def test_no_alloc_direct_gravity_matches_the_allocating_call():
    torch = pytest.importorskip("torch")
    lib = isph.load_library()
    for n, seed in ((6000, 0x7EE52707), (1500, 0x51ED2707)):
        pos, mass = xorshift_debris(n, seed)
        print(f"no-alloc direct gravity: n={n}", flush=True)
        got = _noalloc(lib, torch, pos, mass)
        with isph.GpuSph(n) as s:
            ref = s.gravity_accelerations(pos, mass, SOFT, isph.GRAVITY_DIRECT)
        exact = oracle.direct_gravity(pos, mass, SOFT)
        assert np.isfinite(got).all()
        e_ref = rel_err_floor(got, ref).max()
        e_exact = rel_err_floor(got, exact).max()
        print(f"n={n}: vs allocating call {e_ref:.2e}, vs f64 sum {e_exact:.2e}")
        assert e_ref < TOL and e_exact < TOL


def test_no_alloc_rejects_oversize_and_null():
    lib = isph.load_library()
    fake = ctypes.c_void_p(16)  # never dereferenced: both checks come before any device work
    assert lib.isph_gravity_direct_noalloc(fake, fake, (1 << 20) + 1, SOFT, fake, None) == -34
    assert "exceeds the scratch capacity" in lib.isph_last_error(None).decode()
    assert lib.isph_gravity_direct_noalloc(None, fake, 10, SOFT, fake, None) == isph.EINVAL
