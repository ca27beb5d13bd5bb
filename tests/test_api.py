"""Host-side argument checking of the drop-in API (runs without a GPU: the
checks fire before any device work)."""

import pytest
import torch

import jaxsplat_amd
from jaxsplat_amd import ShapeDtypeError


def _args(n=4):
    return dict(means3d=torch.zeros(n, 3), scales=torch.ones(n, 3), quats=torch.ones(n, 4),
                colors=torch.zeros(n, 3), opacities=torch.ones(n, 1))


KW = dict(viewmat=torch.eye(4), background=torch.zeros(3), img_shape=(16, 16), f=(16.0, 16.0),
          c=(8.0, 8.0), glob_scale=1.0, clip_thresh=0.01)


def test_error_type_is_assertion_and_type_error():
    assert issubclass(ShapeDtypeError, AssertionError)
    assert issubclass(ShapeDtypeError, TypeError)


@pytest.mark.parametrize("name,bad", [("means3d", torch.zeros(4, 2)),
                                      ("scales", torch.zeros(5, 3)),
                                      ("quats", torch.zeros(4, 4, dtype=torch.float64))])
def test_project_shape_and_dtype(name, bad):
    a = _args()
    a[name] = bad
    with pytest.raises(ShapeDtypeError):
        jaxsplat_amd.project(a["means3d"], a["scales"], a["quats"], KW["viewmat"],
                             img_shape=(16, 16), f=(16.0, 16.0), c=(8.0, 8.0), glob_scale=1.0,
                             clip_thresh=0.01, block_width=16)


def test_project_requires_gpu_tensors():
    a = _args()
    with pytest.raises(ShapeDtypeError, match="GPU"):
        jaxsplat_amd.project(a["means3d"], a["scales"], a["quats"], KW["viewmat"],
                             img_shape=(16, 16), f=(16.0, 16.0), c=(8.0, 8.0), glob_scale=1.0,
                             clip_thresh=0.01, block_width=16)


def test_rasterize_shapes():
    n = 4
    with pytest.raises(ShapeDtypeError):
        jaxsplat_amd.rasterize(torch.zeros(n, 3), torch.zeros(n, 1), torch.zeros(3),
                               torch.zeros(n, 2), torch.zeros(n, 1),
                               torch.zeros(n, 1, dtype=torch.int64), torch.zeros(n, 3),
                               torch.zeros(n, 1, dtype=torch.int32), img_shape=(8, 8),
                               block_width=16)


def test_render_rejects_unknown_adjoint():
    with pytest.raises(ValueError):
        jaxsplat_amd.render(**_args(), **KW, adjoint="exact")


@pytest.mark.parametrize("nch,isect,staged", [(3, 0, True), (3, 8_000_000, True),
                                              (1, 2**30 - 1, True), (3, 2**30, False),
                                              (3, 2**31 - 1, False), (4, 1000, False),
                                              (7, 1000, False)])
def test_rasterize_backward_dispatch(nch, isect, staged, monkeypatch):
    """The rasterize backward takes the staged entry point for up to three
    channels and fewer than 2^30 intersections (its record slots are 31-bit
    and it rejects more); everything else goes to the atomic N-channel
    backward, which has neither limit."""
    from jaxsplat_amd import _ops
    monkeypatch.setattr(_ops, "STAGED_BACKWARD", True)
    assert _ops.staged_backward_applies(nch, isect) is staged
    monkeypatch.setattr(_ops, "STAGED_BACKWARD", False)
    assert _ops.staged_backward_applies(nch, isect) is False
