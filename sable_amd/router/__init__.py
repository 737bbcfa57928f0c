"""Learned block router: replaces the reference's scikit-learn block model
(models/density_threshold_spmv.pkl, consumed at
utils/convert_real_to_vbr.py:19-20,58).

* ``features``  per-block features: rows, cols, nnz, density, row-length
                variance, RHS width N;
* ``labels``    per-block GPU cost of every route, measured by the kernels
                themselves (per-tile wall clocks inside the persistent kernel)
                on seeded synthetic blocks;
* ``mlp``       a 2-layer MLP (FP32) trained and evaluated on the GPU by the
                library's forward / backward / Adam kernels;
* ``Router``    picks the route with the lowest predicted cost per block; the
                fixed-threshold plan stays the fallback and the judge: a
                learned plan is kept only if it measures faster at the same
                tolerance (``choose_plan``).
"""
from .features import block_features, FEATURE_NAMES  # noqa: F401
from .router import Router, choose_plan, ROUTES  # noqa: F401
