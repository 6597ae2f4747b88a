"""jaxns_amd -- the jaxns nested-sampling hot path on AMD Instinct MI355X.

Drop-in surface for jaxns' sampling path:
    from jaxns_amd import Model, Prior, NestedSampler, DefaultNestedSampler, TerminationCondition
    from jaxns_amd import distributions as tfpd
    from jaxns_amd.likelihoods import MultivariateNormalLikelihood

The sampling loop (slice chains, batched likelihoods, live-point replacement,
evidence register) and to_results' tree count and evidence scan run as HIP
kernels from libhipns.so (jaxns_amd/csrc); Python drives them through ctypes.
"""
from . import distributions
from .framework import InvalidPriorName, Model, Prior
from .likelihoods import (GaussianMixtureLikelihood, LinearRegressionLikelihood, MultivariateNormalLikelihood,
                          RegressionData)
from .nested_sampler import DefaultNestedSampler, NestedSampler, NestedSamplerState
from .types import (NestedSamplerResults, TerminationCondition, TerminationConditionConjunction,
                    TerminationConditionDisjunction)
from .samplers import UniDimSliceSampler, UniformSampler
from .utils import (load_pytree, load_results, maximum_a_posteriori_point, resample, sample_evidence, save_pytree,
                    save_results, summary)

__all__ = ["Model", "Prior", "NestedSampler", "DefaultNestedSampler", "NestedSamplerState", "TerminationCondition",
           "TerminationConditionConjunction", "TerminationConditionDisjunction", "NestedSamplerResults",
           "MultivariateNormalLikelihood", "GaussianMixtureLikelihood", "LinearRegressionLikelihood", "RegressionData", "summary",
           "save_results", "load_results", "save_pytree", "load_pytree", "resample", "maximum_a_posteriori_point",
           "sample_evidence", "UniDimSliceSampler", "UniformSampler", "distributions", "InvalidPriorName"]
