"""Result and condition types with the field names of jaxns.

TerminationCondition mirrors nested_samplers/common/types.py:26-74 (fields,
``&`` / ``|`` combinators); NestedSamplerResults mirrors types.py:77-98.
"""
from __future__ import annotations

from typing import Any, List, NamedTuple, Optional, Union

import numpy as np


class TerminationCondition(NamedTuple):
    ess: Optional[float] = None
    evidence_uncert: Optional[float] = None
    live_evidence_frac: Optional[float] = None  # deprecated, use dlogZ
    dlogZ: Optional[float] = None
    max_samples: Optional[float] = None
    max_num_likelihood_evaluations: Optional[float] = None
    log_L_contour: Optional[float] = None
    efficiency_threshold: Optional[float] = None
    rtol: Optional[float] = None
    atol: Optional[float] = None
    peak_XL_frac: Optional[float] = None

    def __and__(self, other):
        return TerminationConditionConjunction(conds=[self, other])

    def __or__(self, other):
        return TerminationConditionDisjunction(conds=[self, other])


class TerminationConditionConjunction(NamedTuple):
    conds: List[Any]

    def __and__(self, other):
        return TerminationConditionConjunction(conds=[self, other])

    def __or__(self, other):
        return TerminationConditionDisjunction(conds=[self, other])


class TerminationConditionDisjunction(NamedTuple):
    conds: List[Any]

    def __and__(self, other):
        return TerminationConditionConjunction(conds=[self, other])

    def __or__(self, other):
        return TerminationConditionDisjunction(conds=[self, other])


TERMINATION_BITS = {
    0: "used maximum allowed number of samples",
    1: "evidence uncert below threshold",
    2: "live points evidence below threshold",
    3: "effective sample size big enough",
    4: "used maximum allowed number of likelihood evaluations",
    5: "maximum log-likelihood contour reached",
    6: "sampler efficiency too low",
    7: "entire live-points set is a single plateau",
    8: "relative spread of live points < rtol",
    9: "absolute spread of live points < atol",
    10: "no seed points left",
    11: "XL < max(XL) * peak_XL_frac",
}


class NestedSamplerResults(NamedTuple):
    log_Z_mean: float
    log_Z_uncert: float
    ESS: float
    H_mean: float
    samples: Any
    parametrised_samples: Any
    U_samples: np.ndarray
    log_L_samples: np.ndarray
    log_dp_mean: np.ndarray
    log_X_mean: np.ndarray
    log_posterior_density: np.ndarray
    num_live_points_per_sample: np.ndarray
    num_likelihood_evaluations_per_sample: np.ndarray
    total_num_samples: int
    total_phantom_samples: int
    total_num_likelihood_evaluations: int
    log_efficiency: float
    termination_reason: int


class EvidenceCalculation(NamedTuple):
    log_L: float
    log_X_mean: float
    log_X2_mean: float
    log_Z_mean: float
    log_ZX_mean: float
    log_Z2_mean: float
    log_dZ_mean: float
    log_dZ2_mean: float


class TerminationRegister(NamedTuple):
    num_samples_used: int
    evidence_calc: EvidenceCalculation
    evidence_calc_with_remaining: EvidenceCalculation
    num_likelihood_evaluations: int
    log_L_contour: float
    efficiency: float
    plateau: bool
    no_seed_points: bool
    relative_spread: float
    absolute_spread: float
    peak_log_XL: float


class SampleCollection(NamedTuple):
    """nested_samplers/common/types.py:115-120 (host copies of the device collection)."""
    sender_node_idx: np.ndarray
    log_L: np.ndarray
    U_samples: np.ndarray
    num_likelihood_evaluations: np.ndarray
    phantom: np.ndarray


TermCondType = Union[TerminationCondition, TerminationConditionConjunction, TerminationConditionDisjunction]
