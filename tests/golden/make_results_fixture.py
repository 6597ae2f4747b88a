"""Writes tests/golden/jaxns_results_schema.json: a NestedSamplerResults file laid out as jaxns 2.6.9
writes it (utils.py:588-613 -> internals/namedtuple_utils.py:34-103), built here field by field from
that schema without this package's writer: the NamedTuple as {"type": "__namedtuple__", "__class__":
"jaxns.nested_samplers.common.types.NestedSamplerResults", "__data__": ...}, every array (jax arrays
in a jaxns run, 0-d for the scalars) as {"type": "__jax_ndarray__", "__dtype__", "__data__": base64 of
the raw little-endian bytes, "__shape__"}, and `samples` / `parametrised_samples` as plain dicts of
arrays.  The values are a small synthetic run (M = 6 samples of a 2-D variable "x").

  python tests/golden/make_results_fixture.py
"""
import base64
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def jax_array(a):
    a = np.array(a, copy=True, order="C")  # 0-d arrays stay 0-d
    return {"type": "__jax_ndarray__", "__dtype__": str(a.dtype), "__data__": base64.b64encode(a.tobytes()).decode(),
            "__shape__": list(a.shape)}


def values():
    M = 6
    rng = np.random.default_rng(2024)
    U = rng.uniform(size=(M, 2))
    log_L = np.sort(rng.normal(size=M)) * 2.0
    log_L[0] = -np.inf
    return dict(
        log_Z_mean=np.float64(-3.25), log_Z_uncert=np.float64(0.125), ESS=np.float64(4.5), H_mean=np.float64(-1.75),
        samples={"x": (U * 20.0 - 10.0)}, parametrised_samples={}, U_samples=U, log_L_samples=log_L,
        log_dp_mean=np.array([-np.inf, -3.0, -2.5, -1.5, -1.25, -1.0]),
        log_X_mean=-np.arange(1, M + 1) / 3.0, log_posterior_density=log_L - np.log(400.0),
        num_live_points_per_sample=np.array([4, 4, 4, 3, 2, 1], np.int64),
        num_likelihood_evaluations_per_sample=np.array([1, 1, 3, 7, 9, 4], np.int64),
        total_num_samples=np.int64(M), total_phantom_samples=np.int64(0),
        total_num_likelihood_evaluations=np.int64(25), log_efficiency=np.float64(np.log(M / 25.0)),
        termination_reason=np.int64(4))


def main():
    data = {}
    for k, v in values().items():
        data[k] = {kk: jax_array(vv) for kk, vv in v.items()} if isinstance(v, dict) else jax_array(v)
    doc = {"type": "__namedtuple__", "__class__": "jaxns.nested_samplers.common.types.NestedSamplerResults",
           "__data__": data}
    with open(os.path.join(HERE, "jaxns_results_schema.json"), "w") as f:
        json.dump(doc, f, indent=2)


if __name__ == "__main__":
    main()
