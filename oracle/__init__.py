"""CPU oracle for the attenuation-corrected XRF projector and its solvers.

TEST INFRASTRUCTURE ONLY. Nothing in the product package imports this
package: only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker (or the timed CPU
baseline), never as the thing measured on the GPU.

What it restates (reference citations are relative to the corrct tree):

* ``skimage_ops``  -- scikit-image 0.18.3 ``radon`` (circle=True, square
  image), ``iradon(filter_name=None, circle=False)`` and ``rotate(order=1,
  clip=False)``, as called by ``_projector_backends.py:68,73,333,355``.
  scikit-image is a third-party dependency of corrct (unpinned upstream,
  ``pyproject.toml:35``); version 0.18.3 is the one restated.
* ``attenuation``  -- ``compute_attenuation`` (``_projector_backends.py:43-99``)
  and ``AttenuationVolume.compute_maps`` (``physics/attenuation.py:105-180``).
* ``xrf``          -- ``ProjectorAttenuationXRF.fp/bp`` (``projectors.py:502-622``)
  plus the exact adjoint of the fp (no reference counterpart).
* ``solvers``      -- ``compute_diagonal_scaling``, ``SIRT``, ``MLEM``, ``PDHG`` with
  the l2 / KL data terms and the TV2D regularizer (``solvers.py:60-117,
  688-1206``; ``data_terms.py:217-331, 535-625, 677-757``;
  ``regularizers.py:120-321``; ``operators.py:767-844``).
* ``data_terms``   -- the data dual step and ``compute_residual`` of the l2, wl2,
  l2b, Huber, l1, l12, l1b and KL data terms (``data_terms.py:25-757``).

Every function has an FP32 mode (the arithmetic the reference performs,
operation by operation) and, where it matters, an FP64 mode (same formulas
in double precision) used to state the FP32 error.

How each function is pinned (beyond retyping its formula):

* radon / iradon / rotate: bit-level golden vectors produced by running
  scikit-image 0.18.3 itself (``tests/golden/make_golden_skimage.py``), plus
  closed forms (theta = 0 identity, impulse hats, theta = 90/180 flips).
* compute_attenuation / compute_maps: golden vectors made with scikit-image's
  ``rotate`` composed as the reference composes it, plus mu = 0 => 1,
  multiplicativity, linearity of the log-map in mu, and the theta = 0
  trapezoid closed form.
* XRF fp/bp: reduce to radon/iradon for unit maps and weights; linearity;
  dense explicit matrices at N <= 16 whose transpose must equal the adjoint.
* solvers: a dense-matrix restatement (independent code path) and the
  reference's own convergence tests (``tests/test_solvers.py:100-164``).
* data_terms: each dual step equals the proximal map of sigma times the term's
  convex conjugate found by brute-force 1-D minimisation; special cases that
  reduce one term to another; the group projection's ball property.
"""
