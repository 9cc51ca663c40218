"""The far adjoint's work partition (site ranges of equal estimated cost per workgroup of
four tiles, made on the GPU once per geometry and second radius).

The partition only groups the same terms differently (FP32 runs of 8 sites within a chunk,
FP64 partial per chunk), so any chunk count gives the mixed result within its stated error of
the FP64 adjoint and within rounding of every other count; it is rebuilt when the second
radius changes, so a context that changed it equals, bit for bit, a fresh one made with the
new value."""
import numpy as np
import pytest

from jif3d_amd import _native as nat
from jif3d_amd import synthetic

pytestmark = pytest.mark.gpu
K = nat.KIND_SCALAR_TENSOR


def rel_inf(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def fresh(m, d, monkeypatch, waves=None):
    if waves is None:
        monkeypatch.delenv("JIF3D_AMD_FAR_ADJOINT_WAVES", raising=False)
    else:
        monkeypatch.setenv("JIF3D_AMD_FAR_ADJOINT_WAVES", str(waves))
    ctx = nat.Context(0)
    ctx.configure(m, d)
    ctx.set_mixed_check(0, 0, 5e-7)  # the expansion itself, no fallback
    return ctx


@pytest.mark.parametrize("shape", [(48, 24, 24), (40, 20, 17)])
def test_any_chunk_count_gives_the_same_gradient_within_rounding(monkeypatch, shape):
    n, nzc, nsite = shape
    m, d = synthetic.config4_gravity(n=n, nzc=nzc, nsite=nsite)
    mv = m.GetModelParameters()
    r = np.random.default_rng(21).normal(size=d.nsites * 10) * 1e-9
    g64 = fresh(m, d, monkeypatch).adjoint(K, nat.FP64, mv, r)
    grads = {}
    # 1 wave per tile (one chunk), the default, and one chunk per site or so
    for waves in (1, None, 64 * d.nsites):
        grads[waves] = fresh(m, d, monkeypatch, waves).adjoint(K, nat.MIXED, mv, r)
        assert rel_inf(grads[waves], g64) <= 1e-6, waves
    base = grads[None]
    for waves, g in grads.items():
        assert rel_inf(g, base) <= 1e-9, waves
    # the same context again: the same partition, the same bits
    ctx = fresh(m, d, monkeypatch)
    assert np.array_equal(ctx.adjoint(K, nat.MIXED, mv, r), ctx.adjoint(K, nat.MIXED, mv, r))


def test_partition_follows_the_second_radius(monkeypatch):
    m, d = synthetic.config4_gravity(n=40, nzc=20, nsite=20)
    mv = m.GetModelParameters()
    r = np.random.default_rng(22).normal(size=d.nsites * 10) * 1e-9
    ctx = fresh(m, d, monkeypatch)
    ctx.set_mixed_far_radius(12.0)
    g12 = ctx.adjoint(K, nat.MIXED, mv, r)
    ctx.set_mixed_far_radius(40.0)
    g40 = ctx.adjoint(K, nat.MIXED, mv, r)
    other = fresh(m, d, monkeypatch)
    other.set_mixed_far_radius(40.0)
    assert np.array_equal(g40, other.adjoint(K, nat.MIXED, mv, r))
    other.set_mixed_far_radius(12.0)
    assert np.array_equal(g12, other.adjoint(K, nat.MIXED, mv, r))
