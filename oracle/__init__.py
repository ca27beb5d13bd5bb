"""CPU oracle for the Gaussian-splatting render path (TEST INFRASTRUCTURE ONLY).

This package restates, in NumPy, the semantics of the reference kernels of
yklcs/jaxsplat (``lib/kernels/forward.cu``, ``lib/kernels/backward.cu``,
``lib/kernels/helpers.h``, ``lib/ops.cu``).  It exists only to check the HIP
implementation in ``jaxsplat_amd``:

* ``tests/`` use it as the checker,
* ``__graft_entry__.smoke()`` uses it to check one small GPU invocation,
* ``bench.py`` times it as the CPU baseline (``cpu_baseline`` leg).

Nothing in the product path (``jaxsplat_amd``) imports it, and it must never
be used as a fallback for the GPU path.

Modules
-------
``project``    per-Gaussian projection forward (float64 or a float32
               restatement with the GPU's exact operation order) and backward
               (reference semantics or the consistent true adjoint).
``binning``    tile rectangles, intersection keys, the stable sort order and
               tile bin edges (pure integer logic on the float32 projection).
``rasterize``  per-tile vectorised front-to-back blend and its backward, plus
               brute-force per-pixel loops for tiny inputs.

Pinning (how each function is checked; see ``tests/test_oracle_*.py``)
----------------------------------------------------------------------
* project fwd: closed forms (axis-aligned Gaussian on the optical axis,
  rotation invariance of cov3d), float32 restatement vs float64.
* project bwd, rasterize bwd (consistent mode): torch float64 autograd of an
  independent torch forward, and central finite differences of the float64
  forward.  Reference mode equals consistent mode where no quirk triggers and
  equals the autograd of the unclamped-Jacobian forward for the fov quirk.
* binning: brute force over Python tuples (``sorted``) and ``np.lexsort``.
* rasterize fwd: brute-force sequential per-pixel loop and closed forms
  (single Gaussian, two Gaussians, full occlusion / termination).
"""

from . import project, binning, rasterize  # noqa: F401
