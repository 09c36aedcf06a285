"""MI355X-native pyramid Lucas-Kanade registration (drop-in for getMotion / correctMotion
and the harness around them).

``option`` is the shared parameter dict of the reference's ``utils`` package
(src/wholistic_registration/utils/__init__.py:79-95), with the same defaults;
the harness (``registration.wbi_registration_3d``) writes the solver keys into
it before every ``getMotion`` call, as the reference does.
"""

option = {
    "layer": 3,
    "iter": 10,
    "r": 5,
    "zRatio": 27.693,
    "mask_ref": 0,
    "mask_mov": 0,
    "save_ite": 100,
    "thresFactor": 5,
    "mask_size_range": [5, 500],
    "smoothPenalty_raw": 0.05,
    "tol": 1e-4,
    "expand_radius_xy": 2.0,
    "sigma_grad": 1.0,
    "intensity_k": 2.0,
    "quantile_threshold_q": 0.8,
}
