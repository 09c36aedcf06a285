"""CPU oracle for the pyramid Lucas-Kanade registration path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
directory.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use it, and only as the checker or as
the timed CPU baseline, never as the thing that is measured or shipped.

Contents
--------
``lk_oracle``
    numpy/scipy restatement of ``getMotion`` / ``correctMotion`` and the
    primitives they call (``src/wholistic_registration/utils/
    calFlow3d_Wei_v1.py:37-706``, ``imresize.py:41-191``, ``interp.py:28-45,
    96-100``, ``calculate.py:30-92``, ``preprocess.py:42-66``).  Two modes:

    * ``"ref"``  reproduces the reference's float32/float64 cast points, so it
      matches the reference bit for bit where the reference calls the same
      numpy/scipy routines (pinned by the golden fixtures in
      ``tests/golden``);
    * ``"fp64"`` carries every intensity, coordinate and motion value in
      float64; it is the FP64 oracle the endpoint-error gate is quoted
      against.

``xres_oracle``
    numpy/scipy restatement of the cross-resolution solver ``getMotion_v2``
    with its wrong-region correction (``calFlowCrossResolution.py:
    463-477,1144-2663``), making the reference's own numpy/scipy calls with
    its cast points; pinned bit for bit by fixtures written by running the
    reference (``tests/golden/make_golden_xres.py``) and by the known answers
    of the reference's geometry regression tests.

``metrics``
    endpoint-error statistics restated from
    ``src/wholistic_registration/utils/simulation.py:303-339``.

Parity pinning (what checks each oracle function; see tests/test_oracle_*.py)
---------------------------------------------------------------------------
* ``imresize`` / ``resample_tables``: golden vectors written by importing the
  reference's ``imresize`` (tests/golden/make_golden.py), plus closed forms
  (constant in -> constant out, identity at scale 1, rows sum to one,
  mirror symmetry).
* ``warp_linear``: closed-form trilinear with clamp; identity warp is exact;
  integer shifts are exact.
* ``box3`` / ``patch_sums``: brute force on tiny inputs and a scipy
  ``correlate`` call on the full grid.
* ``nei_diff``: brute force.
* ``solve_patches``: ``numpy.linalg.solve`` on random SPD systems.
* ``spatial_gradient``: constant gradient of linear ramps, halving at faces.
* ``get_motion`` (whole iteration schedule, stop rule and final field):
  golden end-to-end fixtures written by importing the reference.
* ``smooth_penalty_factor``: analytic ramps.
"""
